// pstm_tile.h -- gfx950 Kirchhoff PSTM tile kernels (forward, adjoint,
// velocity gradient), included by the per-mode translation units.
//
// Work decomposition (one workgroup = 4 wave64 = 256 threads):
//   * a tile of TX x TY image columns and 256 consecutive output times;
//     thread t owns output time tau_t for ALL columns of the tile, so the
//     per-column accumulators live in registers and the image is written
//     once per tile, coalesced along tau, without atomics;
//   * the traces that can reach the tile are streamed from the plan's
//     spatial cells (a few contiguous ranges of the sorted trace list),
//     filtered against the tile's aperture/angle reach, and compacted in
//     order (deterministic summation order) into an LDS queue;
//   * each queued trace gets a sample window bounded rigorously from the DSR
//     traveltime over the tile's columns and times; sub-batches of windows
//     are staged HBM -> LDS with coalesced loads (FP32/BF16/FP16 converted to
//     FP32 with the per-trace scale applied), integrated twice in FP64 in LDS
//     when the triangle antialias filter is on;
//   * per (trace, column) the source/receiver horizontal distances are
//     computed once into LDS; each lane then evaluates the DSR time, masks,
//     weight and interpolation for its own output time.  The common case
//     (FULL weight, no taper, clamped record end) is compiled branch-free
//     across the tile's columns so their dependency chains interleave.
//
// Reference method: processors/migration/kirchhoff_kernel.py:459-854
// (migrate_kirchhoff_full) -- same masks (:635, :670-685), DSR time (:688),
// weight (:691-693), clamped linear interpolation (:708-721), image and fold
// sums (:724-730).  The traveltime is evaluated in the equivalent
// time form t_leg = sqrt(tau^2/4 + h^2 / v^2), which is r/v with
// z = v tau / 2 (migration_engine.py:305-306).
#pragma once

#include <type_traits>

#include "pstm_internal.h"

namespace pstm {

enum { MODE_FWD = 0, MODE_ADJ = 1, MODE_GRAD = 2 };

__device__ __forceinline__ float wave_max(float v) {
  for (int off = 32; off > 0; off >>= 1) v = fmaxf(v, __shfl_xor(v, off));
  return v;
}
__device__ __forceinline__ float wave_min(float v) {
  for (int off = 32; off > 0; off >>= 1) v = fminf(v, __shfl_xor(v, off));
  return v;
}
__device__ __forceinline__ float uniform_f(float v) {
  return __int_as_float(__builtin_amdgcn_readfirstlane(__float_as_int(v)));
}
__device__ __forceinline__ int uniform_i(int v) { return __builtin_amdgcn_readfirstlane(v); }

// DPP lane moves (gfx9 encodings).  A lane whose source is outside its row,
// or whose row is masked off, receives 0.
enum : int {
  kDppQuadXor1 = 0xB1,  // quad_perm [1, 0, 3, 2]
  kDppQuadXor2 = 0x4E,  // quad_perm [2, 3, 0, 1]
  kDppRowShr1 = 0x111,
  kDppRowShr2 = 0x112,
  kDppRowShr4 = 0x114,
  kDppRowShr8 = 0x118,
  kDppRowBcast15 = 0x142,  // lane 15 of each row to the next row
  kDppRowBcast31 = 0x143,  // lane 31 to rows 2 and 3
};

template <int CTRL, int ROWS = 0xf>
__device__ __forceinline__ int dpp_move(int v) {
  return __builtin_amdgcn_update_dpp(0, v, CTRL, ROWS, 0xf, false);
}
template <int CTRL, int ROWS = 0xf>
__device__ __forceinline__ float dpp_move(float v) {
  return __int_as_float(dpp_move<CTRL, ROWS>(__float_as_int(v)));
}
template <int CTRL, int ROWS = 0xf>
__device__ __forceinline__ double dpp_move(double v) {
  const long long b = __double_as_longlong(v);
  const int lo = dpp_move<CTRL, ROWS>((int)(b & 0xffffffffll)), hi = dpp_move<CTRL, ROWS>((int)(b >> 32));
  return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
}

// Inclusive prefix sum over the 64 lanes of a wave (row-local doubling
// steps, then the row carries).
template <typename T>
__device__ __forceinline__ T wave_scan(T v, int lane) {
  (void)lane;
  v += dpp_move<kDppRowShr1>(v);
  v += dpp_move<kDppRowShr2>(v);
  v += dpp_move<kDppRowShr4>(v);
  v += dpp_move<kDppRowShr8>(v);
  v += dpp_move<kDppRowBcast15, 0xa>(v);
  v += dpp_move<kDppRowBcast31, 0xc>(v);
  return v;
}

// Value of lane 63.
__device__ __forceinline__ double lane63(double v) {
  const long long b = __double_as_longlong(v);
  const int lo = __builtin_amdgcn_readlane((int)(b & 0xffffffffll), 63);
  const int hi = __builtin_amdgcn_readlane((int)(b >> 32), 63);
  return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
}

// OR over each aligned group of NC (4, 8, 16 or 32) lanes; the result is
// valid in the group's last lane.
template <int NC>
__device__ __forceinline__ int group_or(int v) {
  static_assert(NC == 4 || NC == 8 || NC == 16 || NC == 32, "group of 4, 8, 16 or 32 lanes");
  v |= dpp_move<kDppQuadXor1>(v);
  v |= dpp_move<kDppQuadXor2>(v);
  if constexpr (NC >= 8) v |= dpp_move<kDppRowShr4>(v);
  if constexpr (NC >= 16) v |= dpp_move<kDppRowShr8>(v);
  if constexpr (NC >= 32) v |= dpp_move<kDppRowBcast15, 0xa>(v);
  return v;
}

// Causal then anticausal running sum of one staged FP64 window (one wave):
// A[s] <- sum_{k >= s} sum_{j <= k} A[j].  Each lane sums a contiguous
// segment serially; one wave scan per pass joins the segments.  With C
// non-null the causal sums are also stored, rounded to FP32, in C.  The
// double integral of an FP32 trace is held in FP64 because the triangle taps
// take its second difference; the same symmetric operator is its own
// transpose (adjoint windows).
// SA, SC: element strides of A (doubles) and C (floats).
template <int SA = 1, int SC = 1>
__device__ __forceinline__ void integrate_window(double* A, float* C, int len, int lane) {
  const int seg = ((len + 63) >> 6) | 1;  // odd stride: conflict-free segment reads
  const int s0 = min(lane * seg, len), s1 = min(s0 + seg, len);
  double t = 0.0;
  for (int s = s0; s < s1; ++s) t += A[SA * s];
  double run = wave_scan(t, lane) - t;  // sum of the segments before this lane's
  for (int s = s0; s < s1; ++s) {
    run += A[SA * s];
    A[SA * s] = run;
    if (C) C[SC * s] = (float)run;
  }
  t = 0.0;
  for (int s = s0; s < s1; ++s) t += A[SA * s];
  const double incl = wave_scan(t, lane);
  run = lane63(incl) - incl;  // sum of the segments after this lane's
  for (int s = s1 - 1; s >= s0; --s) {
    run += A[SA * s];
    A[SA * s] = run;
  }
}

// d rounded to the nearest integer as a two's-complement 64-bit value
// (|d| < 2^62): high word floor(d / 2^32), low word the exact remainder.
__device__ __forceinline__ unsigned long long to_fixed64(double d) {
  d = rint(d);
  const double hi = floor(d * 0x1p-32);
  const double lo = fma(hi, -0x1p32, d);  // exact, in [0, 2^32)
  const unsigned long long h = (unsigned long long)(unsigned int)(int)hi;
  return (h << 32) | (unsigned long long)(unsigned int)lo;
}
// Window element of trace sample i in a bank-skewed adjoint window: slot
// i + floor(i / 32) (one spare slot before every 32nd sample).  Its LDS byte
// address base + 8 i + 8 floor(i / 32) is formed with two shift-adds (left to
// itself the compiler spends two shifts and a three-operand add).
typedef __attribute__((address_space(3))) unsigned long long lds_u64;
__device__ __forceinline__ lds_u64* skewed(unsigned long long* base, int i) {
  const uint32_t b = (uint32_t)(uintptr_t)(lds_u64*)base;
  uint32_t t, a;
  asm("v_lshl_add_u32 %0, %1, 3, %2" : "=v"(t) : "v"(i), "v"(b));
  asm("v_lshl_add_u32 %0, %1, 3, %2" : "=v"(a) : "v"(i >> 5), "v"(t));
  return (lds_u64*)(uintptr_t)a;
}
__device__ __forceinline__ void lds_add(lds_u64* q, unsigned long long v) {
  __hip_atomic_fetch_add(q, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
}
// The same for |d| < 2^51 in two instructions: adding 1.5 * 2^52 leaves
// round(d) + 2^51 in the low 52 bits of the mantissa at a fixed exponent, so
// the sum's bit pattern minus that of 1.5 * 2^52 (low word 0, high word
// 0x43380000) is round(d), rounded to nearest even as rint does.
__device__ __forceinline__ unsigned long long to_fixed64_small(double d) {
  return (unsigned long long)__double_as_longlong(d + 0x1.8p52) - 0x4338000000000000ull;
}
__device__ __forceinline__ double from_fixed64(long long v) {
  return fma((double)(int)(v >> 32), 0x1p32, (double)(unsigned int)(v & 0xffffffffll));
}

template <int DT>
__device__ __forceinline__ float load_sample(const void* tr, int64_t idx) {
  if constexpr (DT == PSTM_F32) {
    return ((const float*)tr)[idx];
  } else if constexpr (DT == PSTM_BF16) {
    uint32_t b = ((const uint16_t*)tr)[idx];
    return __uint_as_float(b << 16);
  } else {
    uint16_t b = ((const uint16_t*)tr)[idx];
    return (float)(*(const _Float16*)&b);
  }
}

// Sample at byte offset `boff` of a raw buffer over one trace (0 at and past
// its end).
template <int DT>
__device__ __forceinline__ float buffer_sample(__amdgpu_buffer_rsrc_t r, int boff) {
  if constexpr (DT == PSTM_F32) {
    return __uint_as_float(__builtin_amdgcn_raw_buffer_load_b32(r, boff, 0, 0));
  } else if constexpr (DT == PSTM_BF16) {
    return __uint_as_float((uint32_t)__builtin_amdgcn_raw_buffer_load_b16(r, boff, 0, 0) << 16);
  } else {
    const uint16_t b = __builtin_amdgcn_raw_buffer_load_b16(r, boff, 0, 0);
    return (float)__builtin_bit_cast(_Float16, b);
  }
}

// Compensated two-term update: (s, c) += y.
// Knuth's branch-free TwoSum recovers the rounding error of s + y exactly
// with six additions (no magnitude test); contraction stays off so that no
// addition is fused with the product forming y.
__device__ __forceinline__ void two_sum_add(float& s, float& c, float y) {
#pragma clang fp contract(off)
  const float t = s + y;
  const float z = t - s;
  c += (s - (t - z)) + (y - z);
  s = t;
}

// Cosine edge taper of x towards its limit xmax (processors/migration/
// aperture.py:226-270, the shape of weights.py:192-236): 1 up to
// s = xmax (1 - tw), 0.5 (1 + cos(pi (x - s) / (xmax - s))) on (s, xmax],
// 0 beyond (selects: a zero-width zone never reaches the division's NaN).
__device__ __forceinline__ float edge_taper(float x, float xmax, float tw) {
  const float s = xmax * (1.0f - tw);
  const float c = 0.5f * (1.0f + __cosf(3.14159265358979f * __fdividef(x - s, xmax - s)));
  return x > xmax ? 0.f : (x > s ? c : 1.f);
}

// Taper composition of processors/migration/aperture.py:147-182: distance
// tapers of both legs towards the aperture (time-dependent under
// PSTM_FLAG_TIME_APERTURE) times angle tapers of both legs, angle
// atan2(h, z).  hs2, hr2: squared leg distances (m^2); z = v tau / 2 (m).
// The trace's offset taper multiplies its staged samples (forward, gradient)
// or its flushed window (adjoint) instead.
__device__ __forceinline__ float aperture_taper(float hs2, float hr2, float z, const TileArgs& a) {
  float ap = a.aperture;
  if (a.flags & PSTM_FLAG_TIME_APERTURE) ap = fminf(ap, z * a.tan_th);
  const float hs = __builtin_amdgcn_sqrtf(hs2), hr = __builtin_amdgcn_sqrtf(hr2), tw = a.taper_width;
  // each factor is 1 in its flat zone (distance up to ap (1 - tw), angle up
  // to th_max (1 - tw), i.e. h <= z tan_taper): a wave evaluates the cosine
  // (and the angle) only when one of its lanes lies beyond
  const float sd = ap * (1.0f - tw), tz = z * a.tan_taper;
  float w = 1.f;
  if (hs > sd) w *= edge_taper(hs, ap, tw);
  if (hr > sd) w *= edge_taper(hr, ap, tw);
  if (hs > tz) w *= edge_taper(atan2f(hs, z), a.th_max, tw);
  if (hr > tz) w *= edge_taper(atan2f(hr, z), a.th_max, tw);
  return w;
}

// Offset window taper of one trace (aperture.py:184-224): near taper
// 0.5 (1 - cos(pi (o - min) / r)) below min + r, far taper
// 0.5 (1 + cos(pi (o - (max - r)) / r)) above max - r, r = tw (max - min).
// The plan keeps only traces inside [min, max]; 1 without a window.
__device__ __forceinline__ float offset_taper(float off, const TileArgs& a) {
  if (!(a.off_max > 0.f)) return 1.f;
  const float r = a.taper_width * (a.off_max - a.off_min);
  const float pi = 3.14159265358979f;
  const float nr = off < a.off_min + r ? 0.5f * (1.0f - __cosf(pi * (off - a.off_min) / r)) : 1.f;
  const float fr = off > a.off_max - r ? 0.5f * (1.0f + __cosf(pi * (off - (a.off_max - r)) / r)) : 1.f;
  return nr * fr;
}

// GEN = false: FULL weight, no taper, clamped record end (compile-time);
// GEN = true: weight mode, taper and zero-outside read from the arguments.
//
// Inside the kernel traveltimes are in samples: with qd = 1/(v dt)^2 and
// tqs = tau^2/(4 dt^2) a leg is t/dt = sqrt(h^2 qd + tqs) and rs = dt/t.
// Each weight is a per-lane, per-column constant cw, applied once per
// sub-batch, times a power of prs = rs rr = dt^2/(t_s t_r):
//   FULL       tau^2/4 q (t_s t_r)^-2 = cw prs^2,  cw = tau^2/4 q / dt^4
//   SPREADING  q (t_s t_r)^-1         = cw prs,    cw = q / dt^2
//   OBLIQUITY  tau^2/4 (t_s t_r)^-1   = cw prs,    cw = tau^2/4 / dt^2
//   NONE       1                      = cw,        cw = 1
//
// OCC: workgroups per CU the kernel is built for (registers and LDS).
// FIXED (adjoint): the windows accumulate in 64-bit fixed point with integer
// LDS atomics (float LDS atomics run at ~1/3 lane per clock on gfx950,
// 64-bit integer ones at ~9).  Each workgroup scales by 2^s chosen from a
// bound G of its own contributions so that one contribution is < 2^47 units
// and a window sample (<= 2^13 contributions of <= 4 taps) stays < 2^62; the
// accumulation is exact, each contribution is rounded once at 2^-47 G.
// Without antialias (SKEW) the taps are rounded in FP32 to integers below
// 2^30 at a scale 2^(30 - ceil(log2 G)) and the window sample of trace
// sample i sits at i + floor(i / 32) (one spare slot before every 32nd
// sample, which receives the second tap of a pair that straddles the gap and
// is added back at the flush): with pairs of waves interleaved over each half
// of the block's times, lanes sixteen apart would otherwise hit the same LDS
// bank whenever the moveout is one sample per output sample.
//
// V1 (velocity given per output time only, PSTM_VEL_1D): the velocity terms
// are the same for every column of the tile and are held once per lane.
// WM: the weight mode compiled into the fast forms (GEN = false): PSTM_WEIGHT_FULL
// or PSTM_WEIGHT_SPREADING (the class API's default); the general forms read it.
template <int MODE, int TX, int TY, int DT, bool AA, bool VTI, bool GEN, int OCC, bool FIXED, bool V1,
          int WM = PSTM_WEIGHT_FULL, bool CAN_SEG = false>
__global__ __launch_bounds__(kThreads, OCC) void k_tile(TileArgs a) {
  constexpr int NC = TX * TY;
  constexpr bool FX = MODE == MODE_ADJ && FIXED;
  constexpr bool SKEW = FX && !AA;  // FP32 taps rounded to integers, bank-skewed windows
  // Forward fast path: traces that pass the masks at every time and column
  // of a wave and whose tile times never reach past the record are summed
  // without the mask, clamp and fold instructions (their fold is counted once
  // per wave); the rest take the masked loop.  Both give the same sums.
  constexpr bool FULLPATH = MODE == MODE_FWD && !GEN;
  // Adjoint counterpart (plain fixed-point forms): traces that stay inside the
  // record in a wave AND pass every mask at every time and column of it
  // scatter without the mask and without the per-lane skip branch.
  constexpr bool ADJ_UNM = MODE == MODE_ADJ && !GEN && !AA;
  // Software prefetch of the next trace's distances (registers permitting:
  // the plain forward and gradient forms).
  constexpr bool PREFETCH = MODE != MODE_ADJ && !AA && !VTI && !GEN && (OCC <= 3 || (V1 && OCC <= 4));
  static_assert(NC == 8 || (NC == 32 && MODE == MODE_ADJ), "columns per tile: 8 (32 for the adjoint)");
  static_assert(!FIXED || MODE == MODE_ADJ, "fixed-point windows are an adjoint form");
  // ILV (antialias forward at 3 or more workgroups per CU): the FP64 double
  // integral A at an 8-byte stride and the FP32 single integral C in a second
  // array, also at an 8-byte stride, a fixed 8 CAPE bytes after A, so that a
  // tap reads both through one address (C through the instruction's
  // immediate offset).  Neighbouring samples of A fill distinct bank pairs;
  // C's 8-byte stride costs a 2-way conflict on its 4-byte reads (16-byte
  // elements holding A and C side by side conflicted 2-way on A, 4-way on C).
  constexpr bool ILV = AA && MODE == MODE_FWD && OCC >= 3;
  // PAIR (plain forward at 4 or more workgroups per CU): each window element
  // is a sample and its forward difference (a[i], a[i+1] - a[i]) in FP32, so
  // a contribution reads one 8-byte element and interpolates with one FMA
  // (the same FP32 difference the single-sample layout forms per read)
  constexpr bool PAIR = MODE == MODE_FWD && !AA && !GEN && !VTI && OCC >= 4;
  // CAN_SEG: the form launched for traces whose whole window exceeds the
  // LDS share of the lowest-occupancy form; it stages longer windows in
  // segments (a separate instantiation, so the other forms carry none of it).
  static_assert(!CAN_SEG || GEN, "segmenting forms are general forms");
  constexpr int CAPF = sample_cap_floats<NC, AA, OCC>();         // staged-window floats
  constexpr int CAPE = ILV ? CAPF / 4 : PAIR ? CAPF / 2 : window_cap_elems<NC, AA, OCC, FX>();  // staged-window elements
  constexpr int SA = 1;                                          // element stride in doubles (Wa)
  constexpr int SC = ILV ? 2 : 1;                                // element stride in floats (Wc)
  // every staged element, FP64 A and FP32 C, lies inside samp[CAPF]
  static_assert(!AA || (ILV ? 4 * CAPE <= CAPF : 3 * CAPE <= CAPF), "antialias window layout exceeds the LDS buffer");
  constexpr int MB = batch_of(OCC);  // traces per sub-batch
  __shared__ TileShared<MB> S;
  __shared__ PairShared<NC, AA, MB> P;
  __shared__ __align__(16) float samp[CAPF];
  // Plain adjoint: per trace of the sub-batch, the sum of its contributions
  // that read past the record end (all land on samples n-2 and n-1; split
  // between them at the flush).  The MB floats fit in the LDS budget's slack
  // of the forms at 5 or more workgroups per CU; the others scatter a wave sum
  // instead (their LDS shares end within a granule of the per-CU capacity).
  constexpr bool CLAMPSUM = MODE == MODE_ADJ && !AA && OCC >= 5;
  [[maybe_unused]] __shared__ float clampS[CLAMPSUM ? MB : 1];
  // Staged windows: FP32 samples, or with antialias the FP64 double integral
  // (Wa) and the FP32 single integral (Wc, forward only).
  [[maybe_unused]] float* const Wf = samp;
  [[maybe_unused]] float2* const Wp = reinterpret_cast<float2*>(samp);  // PAIR elements
  [[maybe_unused]] double* const Wa = reinterpret_cast<double*>(samp);
  [[maybe_unused]] float* const Wc = samp + 2 * CAPE;  // ILV: stride 2 floats
  [[maybe_unused]] long long* const Wl = reinterpret_cast<long long*>(samp);  // adjoint, fixed point

  const int tid = threadIdx.x, lane = tid & 63, wave = __builtin_amdgcn_readfirstlane(tid >> 6);

  // ---- block -> (time block, tile): deepest time blocks first, each XCD a
  // contiguous spatial chunk of tiles.
  const int n_tiles = a.n_tiles_x * a.n_tiles_y;
  const int chunk = (n_tiles + 7) / 8;
  const int xcd = blockIdx.x & 7, r = blockIdx.x >> 3;
  const int tb = a.n_tblocks - 1 - r / chunk;
  const int tile = xcd * chunk + r % chunk;
  if (tile >= n_tiles || tb < 0) return;
  const int tx = tile / a.n_tiles_y, ty = tile % a.n_tiles_y;

  // ---- per-lane constants
  // The adjoint interleaves pairs of waves over each half of the block's
  // times (waves 2h and 2h + 1 own tau = 128 h + (w & 1) + 2 j): neighbouring
  // lanes are 2 samples apart, so the taps of one (trace, column) land on
  // distinct window samples in one atomic instruction unless the moveout is
  // below 1/2 sample per sample, while a wave spans 128 times (all four waves
  // interleaved over 256 times leave more lanes outside the masks: 1.6 %
  // slower; contiguous times put many lanes' taps on one sample: 7 % slower).
  //
  // The time blocks are aligned to the END of the output times: block tb
  // starts at tb * kTileTau - kshift, so the one ragged block (with idle
  // lanes) is the shallowest, where the angle mask admits the fewest traces,
  // not the deepest and heaviest one.
  const int kshift = a.n_tblocks * kTileTau - a.n_tau;  // 0 .. kTileTau - 1
  const int kb0 = tb * kTileTau - kshift;               // first time of the block (may be < 0)
  const int k = MODE == MODE_ADJ ? kb0 + (wave >> 1) * 128 + lane * 2 + (wave & 1) : kb0 + tid;
  const bool kvalid = k >= 0;  // k < n_tau always
  const int kk = kvalid ? k : 0;
  const float tau = a.tau0 + (float)kk * a.dtau;
  const float tq = 0.25f * tau * tau;
  const float idt2 = a.inv_dt * a.inv_dt;
  const float tqs = fmaxf(tq, 1e-12f) * idt2;
  [[maybe_unused]] const float tqs2 = tqs * tqs;  // VTI: constant term of the anelliptic numerator
  const int wmode = GEN ? a.weight_mode : WM;
  [[maybe_unused]] const float aa_scale = AA ? a.aa_scale_over_dt / a.inv_dt : 0.f;
  [[maybe_unused]] const float sil = a.d_il * aa_scale, sxl = a.d_xl * aa_scale;  // antialias: scaled column steps

  float qd[NC], hl2[NC];
  [[maybe_unused]] float k1[VTI ? NC : 1], c1[VTI ? NC : 1];  // 1 + 2 eta, 2 (1 + eta) tqs
  [[maybe_unused]] float sv[MODE == MODE_GRAD ? NC : 1];  // 1/v
  // eta extremes of the tile (VTI), each clamped to 0 from its side: they
  // bound the anelliptic leg time between sqrt(t0^2 + p / (1 + 2 max(eta, 0)))
  // and sqrt(t0^2 + p / (1 + 2 min(eta, 0))), p = h^2 / v^2 (1 + 2 eta > 0)
  float emax_l = 0.f, emin_l = 0.f;
  float qmin_l = 3.0e38f, qmax_l = 0.f, hmax_l = -1.f;
  unsigned cvalid = 0;  // bit c: column c of the tile is inside the image
#pragma unroll
  for (int c = 0; c < NC; ++c) {
    const int il = tx * TX + c / TY, xl = ty * TY + c % TY;
    const bool cv = il < a.n_il && xl < a.n_xl;
    cvalid |= cv ? (1u << c) : 0u;
    const int64_t vi = (V1 || a.vel_mode == PSTM_VEL_1D)
                           ? (int64_t)kk
                           : ((int64_t)(cv ? il : 0) * a.n_xl + (cv ? xl : 0)) * a.n_tau + kk;
    const float v = a.vel[vi];
    // with V1 the columns outside the image keep the common values (their
    // sums are never written and their adjoint input is zero)
    const bool ok = kvalid && (V1 || cv) && v > 0.f;
    const float vv = ok ? v : 1.f;
    const float q = 1.0f / (vv * vv);
    qd[c] = q * idt2;
    if constexpr (MODE == MODE_GRAD) sv[c] = 1.0f / vv;
    const float z = 0.5f * vv * tau;
    float ap = a.aperture;
    if (a.flags & PSTM_FLAG_TIME_APERTURE) ap = fminf(ap, z * a.tan_th);
    const float hl = fminf(ap, a.tan_th * (z + 1e-6f));
    hl2[c] = ok ? hl * hl : -1.f;
    if constexpr (VTI) {
      const float e = ok ? a.eta[vi] : 0.f;
      k1[c] = 1.0f + 2.0f * e;
      c1[c] = (2.0f + 2.0f * e) * tqs;
      emax_l = fmaxf(emax_l, e);
      emin_l = fminf(emin_l, e);
    }
    if (ok) {
      qmin_l = fminf(qmin_l, q);
      qmax_l = fmaxf(qmax_l, q);
    }
    hmax_l = fmaxf(hmax_l, hl2[c]);
  }
  // Weight constant of column c: tau^2/4 q / dt^4 (FULL), q / dt^2
  // (SPREADING), tau^2/4 / dt^2 (OBLIQUITY) or 1 (NONE), formed from qd when
  // needed (at the flushes) rather than held per column in registers.
  const float cwt = tq * idt2;
  auto cw = [&](int c) -> float {
    if (wmode == PSTM_WEIGHT_FULL) return cwt * qd[c];
    if (wmode == PSTM_WEIGHT_SPREADING) return qd[c];
    if (wmode == PSTM_WEIGHT_OBLIQUITY) return cwt;
    return 1.f;
  };
  // per-wave column reach (max and min over the wave's times)
#pragma unroll
  for (int c = 0; c < NC; ++c) {
    const float hc = wave_max(hl2[c]);
    if (lane == 0) P.hcol[wave][c] = hc;
    if constexpr (FULLPATH || ADJ_UNM) {
      const float hn = wave_min(hl2[c]);
      if (lane == 0) P.hcolmin[wave][c] = hn;
    }
  }
  {
    const float hw = wave_max(hmax_l), qn = wave_min(qmin_l), qx = wave_max(qmax_l);
    [[maybe_unused]] const float em = VTI ? wave_max(emax_l) : 0.f, en = VTI ? wave_min(emin_l) : 0.f;
    if (lane == 0) {
      S.hwave[wave] = hw;
      S.qmin[wave] = qn;
      S.qmax[wave] = qx;
      // first and last valid time of the wave (a wave with none has no reach)
      const int ka = max(MODE == MODE_ADJ ? kb0 + (wave >> 1) * 128 + (wave & 1) : kb0 + wave * 64, 0);
      const int kb = max(MODE == MODE_ADJ ? kb0 + (wave >> 1) * 128 + 126 + (wave & 1) : kb0 + wave * 64 + 63, 0);
      const float ta = a.tau0 + (float)ka * a.dtau, tbv = a.tau0 + (float)kb * a.dtau;
      S.tqa[wave] = 0.25f * ta * ta;
      S.tqb[wave] = 0.25f * tbv * tbv;
      if constexpr (VTI) {
        S.emaxw[wave] = em;
        S.eminw[wave] = en;
      }
      if (wave == 0) S.queue_n = 0;
    }
  }
  __syncthreads();
  float hblk = -1.f;
#pragma unroll
  for (int w = 0; w < kWaves; ++w) hblk = fmaxf(hblk, S.hwave[w]);
  float emax = 0.f, emin = 0.f;
  if constexpr (VTI) {
#pragma unroll
    for (int w = 0; w < kWaves; ++w) {
      emax = fmaxf(emax, S.emaxw[w]);
      emin = fminf(emin, S.eminw[w]);
    }
  }

  // ---- accumulators
  float acc[NC], accb[NC], cmp[NC];
  float fcnt[NC];  // fold counts (exact in FP32 below 2^24)
  int nfull = 0;  // traces summed by the unmasked loop (they count at every column and time of the wave)
  [[maybe_unused]] float m_in[MODE == MODE_ADJ ? NC : 1];
#pragma unroll
  for (int c = 0; c < NC; ++c) {
    acc[c] = 0.f;
    accb[c] = 0.f;
    cmp[c] = 0.f;
    fcnt[c] = 0.f;
  }
  if constexpr (MODE == MODE_ADJ) {
#pragma unroll
    for (int c = 0; c < NC; ++c) {
      const int il = tx * TX + c / TY, xl = ty * TY + c % TY;
      m_in[c] = (kvalid && ((cvalid >> c) & 1u)) ? a.image_in[((int64_t)il * a.n_xl + xl) * a.n_tau + k] * cw(c)
                                                 : 0.f;
    }
  }
  [[maybe_unused]] double fx_inv = 0.0;
  bool narrow = false;  // SKEW: taps in FP32 at 2^(30 - e) (else FP64 at 2^(47 - e))
  if constexpr (FX) {
    // bound of one contribution: |m cw| times the largest weight power
    // (rs, rr <= 1/sqrt(tqs)), times 4 for the antialias taps
    const float wdmax = wmode == PSTM_WEIGHT_FULL ? 1.0f / (tqs * tqs)
                                                  : (wmode == PSTM_WEIGHT_NONE ? 1.0f : 1.0f / tqs);
    float gl = 0.f;
#pragma unroll
    for (int c = 0; c < NC; ++c) gl = fmaxf(gl, fabsf(m_in[c]) * wdmax);
    if (AA) gl *= 4.f;
    const float gw = wave_max(gl), gn = wave_min(gl > 0.f ? gl : 3.0e38f);
    if (lane == 0) {
      S.gmax[wave] = gw;
      S.gmin[wave] = gn;
    }
    __syncthreads();
    float G = 0.f, Gmin = 3.0e38f;
#pragma unroll
    for (int w = 0; w < kWaves; ++w) {
      G = fmaxf(G, S.gmax[w]);
      Gmin = fminf(Gmin, S.gmin[w]);
    }
    if (!(G > 0.f)) return;  // an all-zero image tile scatters nothing (uniform over the workgroup)
    G = fminf(G, 3.0e38f);
    int e;
    (void)frexpf(G, &e);  // G < 2^e
    // FP32 taps rounded at 2^-30 G when the per-lane bounds span at most a
    // factor kNarrowSpan; else FP64 taps rounded at 2^-47 G
    // (the image values carry the scale, a power of two: exact in FP32)
    narrow = SKEW && G <= kNarrowSpan * Gmin;
    const int sh = narrow ? 30 - e : 47 - e;  // taps |w m| 2^sh < 2^30 or 2^47
#pragma unroll
    for (int c = 0; c < NC; ++c) m_in[c] = ldexpf(m_in[c], sh);
    fx_inv = ldexp(1.0, -sh);
  }

  const bool use_taper = GEN ? (a.flags & PSTM_FLAG_TAPER) != 0 : false;
  const bool zero_out = GEN ? (a.flags & PSTM_FLAG_ZERO_OUTSIDE) != 0 : false;

  if (hblk > 0.f) {
    // tile rectangle over valid columns (local coordinates)
    const int il_hi = min(tx * TX + TX - 1, a.n_il - 1), xl_hi = min(ty * TY + TY - 1, a.n_xl - 1);
    const float X0 = (float)(tx * TX) * a.d_il, X1 = (float)il_hi * a.d_il;
    const float Y0 = (float)(ty * TY) * a.d_xl, Y1 = (float)xl_hi * a.d_xl;
    const float R = sqrtf(hblk);
    int cx0 = (int)floorf((X0 - R - a.cell_x0) / a.cell_w);
    int cx1 = (int)floorf((X1 + R - a.cell_x0) / a.cell_w);
    int cy0 = (int)floorf((Y0 - R - a.cell_y0) / a.cell_h);
    int cy1 = (int)floorf((Y1 + R - a.cell_y0) / a.cell_h);
    cx0 = max(cx0, 0);
    cy0 = max(cy0, 0);
    cx1 = min(cx1, a.ncx - 1);
    cy1 = min(cy1, a.ncy - 1);
    const int pad = AA ? a.aa_pad : 0;
    const int wlo = -pad - kWindowMargin, whi = a.n_samples - 1 + pad + kWindowMargin;

    // Partial sums (accb) of at least kFlushTraces traces per wave, summed in
    // FP32, are added to the running sums (acc) between sub-batches, with the
    // compensated update under PSTM_FLAG_KAHAN.
    int nsum = 0;
    auto join_partials = [&]() {
      if (a.flags & PSTM_FLAG_KAHAN) {
#pragma unroll
        for (int c = 0; c < NC; ++c) {
          two_sum_add(acc[c], cmp[c], accb[c] * cw(c));
          accb[c] = 0.f;
        }
      } else {
#pragma unroll
        for (int c = 0; c < NC; ++c) {
          acc[c] = fmaf(accb[c], cw(c), acc[c]);
          accb[c] = 0.f;
        }
      }
    };
    // Process the queued traces in LDS-sized sub-batches.  `final_flush`:
    // also process a last partial batch.
    auto process_queue = [&](bool final_flush) {
      int qs = 0;
      const int qn = S.queue_n;
      // One sub-batch: the nb queued windows from queue position qs, placed
      // at S.b_off / S.b_base.  SEG: a single window that is one segment of a
      // longer one; only the contributions whose sample index lies in
      // [sg_lo, sg_hi) are summed (every contribution of the trace belongs to
      // exactly one segment), and every lane's read position is held inside
      // the staged segment.
      auto run_batch = [&](const int qs, const int nb, const bool seg, const int sg_lo, const int sg_hi,
                           const float sg_rd) {
        (void)sg_lo;
        (void)sg_hi;
        (void)sg_rd;
        // -- stage (FWD/GRAD: load samples; ADJ: zero the windows)
        for (int b = wave; b < nb; b += kWaves) {
          const int e = qs + b;
          const int w0 = S.q_w0[e], len = S.q_len[e], off = S.b_off[b];
          if constexpr (MODE == MODE_ADJ) {
            const int span = SKEW ? skew_len(len) : len;
            if constexpr (CLAMPSUM) {
              if (lane == 0) clampS[b] = 0.f;
            }
            for (int s = lane; s < span; s += 64) {
              if constexpr (FX) Wl[off + s] = 0;
              else if constexpr (AA) Wa[off + s] = 0.0;
              else Wf[off + s] = 0.f;
            }
          } else {
            const int id = a.ids[S.q_j[e]];
            // per-trace factors: the storage scale, and the offset taper (the
            // operator is linear in the trace samples)
            const float sc = (a.scale ? a.scale[id] : 1.f) * (use_taper ? offset_taper(a.offs[S.q_j[e]], a) : 1.f);
            const int64_t row = (int64_t)id * a.n_samples;
            if constexpr (PAIR) {
#pragma clang fp contract(off)
              // A raw buffer over the window and the sample after it (the
              // window lies inside the record): the lanes past the window end
              // fetch nothing (the range check returns 0), and at the record's
              // last sample the difference reads 0 (never interpolated: reads
              // stop at sample n - 2).  Products and difference are rounded as
              // the single-sample layout rounds them.
              constexpr int ES = DT == PSTM_F32 ? 4 : 2;
              const __amdgpu_buffer_rsrc_t rs = __builtin_amdgcn_make_buffer_rsrc(
                  const_cast<char*>(static_cast<const char*>(a.traces)) + (row + w0) * ES, 0,
                  min(len + 1, a.n_samples - w0) * ES, kRawBufferWord3);
              for (int s0 = 0; s0 < len; s0 += 4 * 64) {
                float v[4], vn[4];
#pragma unroll
                for (int u = 0; u < 4; ++u) {
                  const int boff = (s0 + lane + 64 * u) * ES;
                  v[u] = buffer_sample<DT>(rs, boff);
                  vn[u] = buffer_sample<DT>(rs, boff + ES);
                }
#pragma unroll
                for (int u = 0; u < 4; ++u) {
                  const int s = s0 + lane + 64 * u;
                  if (s < len) {
                    const float x = v[u] * sc;
                    Wp[off + s] = make_float2(x, vn[u] * sc - x);
                  }
                }
              }
              continue;
            }
            if constexpr (AA && MODE == MODE_FWD) {
              if (a.int_a) {
                // traces integrated once per call: copy the window of A and C
                // (storage scale already applied; the offset taper scales both),
                // constant before the record, and past its end C = C(n-1) = A(n-1)
                // with A falling linearly: A(n-1+k) = (1 - k) A(n-1)
                const float ot = use_taper ? offset_taper(a.offs[S.q_j[e]], a) : 1.f;
                const double* const ra = a.int_a + row;
                const float* const rc = a.int_c + row;
                const int n = a.n_samples;
                if (w0 >= 0 && w0 + len <= n && !use_taper) {
                  // the window lies inside the record: a plain copy
                  for (int s0 = 0; s0 < len; s0 += 4 * 64) {
                    double av[4];
                    float cv[4];
#pragma unroll
                    for (int u = 0; u < 4; ++u) {
                      const int src = w0 + min(s0 + lane + 64 * u, len - 1);
                      av[u] = ra[src];
                      cv[u] = rc[src];
                    }
#pragma unroll
                    for (int u = 0; u < 4; ++u) {
                      const int s = s0 + lane + 64 * u;
                      if (s < len) {
                        Wa[SA * (off + s)] = av[u];
                        Wc[SC * (off + s)] = cv[u];
                      }
                    }
                  }
                  continue;
                }
                // four independent pairs of loads per lane in flight (indexes
                // clamped into the record), then the extension and the stores
                for (int s0 = 0; s0 < len; s0 += 4 * 64) {
                  double av[4];
                  float cv[4];
#pragma unroll
                  for (int u = 0; u < 4; ++u) {
                    const int src = min(max(w0 + s0 + lane + 64 * u, 0), n - 1);
                    av[u] = ra[src];
                    cv[u] = rc[src];
                  }
#pragma unroll
                  for (int u = 0; u < 4; ++u) {
                    const int s = s0 + lane + 64 * u;
                    const int src = w0 + s;
                    if (s < len) {
                      // past the end: A(n-1+k) = (1 - k) A(n-1), C = A(n-1)
                      const double A = src < n ? av[u] : (double)(n - src) * av[u];
                      const float C = src < 0 ? 0.f : (src < n ? cv[u] : (float)av[u]);
                      Wa[SA * (off + s)] = A * (double)ot;
                      Wc[SC * (off + s)] = C * ot;
                    }
                  }
                }
                continue;
              }
            }
            // in-record part [lo, hi) of the window; antialias windows also
            // hold zero padding outside the record
            const int lo = max(w0, 0), hi = min(w0 + len, a.n_samples);
            if (lo > w0 || hi < w0 + len) {
              for (int s = lane; s < len; s += 64) {
                const int src = w0 + s;
                if (src < lo || src >= hi) {
                  if constexpr (AA) Wa[SA * (off + s)] = 0.0;
                  else Wf[off + s] = 0.f;
                }
              }
            }
            // four independent loads per lane in flight (indexes clamped into
            // the record, so every load is in bounds), then convert and store
            for (int s0 = lo - w0; s0 < hi - w0; s0 += 4 * 64) {
              float v[4];
#pragma unroll
              for (int u = 0; u < 4; ++u) {
                const int src = min(w0 + s0 + lane + 64 * u, hi - 1);
                v[u] = load_sample<DT>(a.traces, row + src);
              }
#pragma unroll
              for (int u = 0; u < 4; ++u) {
                const int s = s0 + lane + 64 * u;
                if (w0 + s < hi) {
                  if constexpr (AA) Wa[SA * (off + s)] = (double)(v[u] * sc);
                  else Wf[off + s] = v[u] * sc;
                }
              }
            }
          }
        }
        // -- (trace, column) distances and the waves each trace reaches.  The
        // NC columns of a trace sit in NC consecutive lanes of one wave.
        for (int idx = tid; idx < nb * NC; idx += kThreads) {
          const int b = idx / NC, c = idx % NC;
          const float4 g = a.geom[S.q_j[qs + b]];
          const float X = (float)(tx * TX + c / TY) * a.d_il, Y = (float)(ty * TY + c % TY) * a.d_xl;
          const float dxs = X - g.x, dys = Y - g.y, dxr = X - g.z, dyr = Y - g.w;
          const float hs = fmaf(dxs, dxs, dys * dys), hr = fmaf(dxr, dxr, dyr * dyr);
          const float hm = fmaxf(hs, hr);
          P.h2[b][c] = make_float2(hs, hr);
          P.hm[b][c] = hm;
          // the triangle half-width's scale is folded in: L = qd (|g.x rs + g.z rr| + |g.y rs + g.w rr|)
          if constexpr (AA) P.g[b][c] = make_float4(dxs * sil, dys * sxl, dxr * sil, dyr * sxl);
          int bits = 0, nfb = 0;
#pragma unroll
          for (int w = 0; w < kWaves; ++w) {
            bits |= hm < P.hcol[w][c] ? (1 << w) : 0;
            if constexpr (FULLPATH || ADJ_UNM) nfb |= hm < P.hcolmin[w][c] ? 0 : (1 << w);
          }
          bits = group_or<NC>(bits);
          if constexpr (FULLPATH || ADJ_UNM) nfb = group_or<NC>(nfb);  // a wave some column does not fully pass
          // bits 4-7: passes every mask and stays in the record; bits 8-11: passes
          // every mask but some time reads past the record (clamped)
          if constexpr (FULLPATH) bits |= ((S.q_nc[qs + b] & ~nfb & 0xf) << 4) | ((~S.q_nc[qs + b] & ~nfb & 0xf) << 8);
          else if constexpr (ADJ_UNM)
            bits |= ((S.q_nc[qs + b] & 0xf) << 4) | ((S.q_nc[qs + b] & ~nfb & 0xf) << 8) |
                    ((~S.q_nc[qs + b] & ~nfb & 0xf) << 12);
          else if constexpr (MODE == MODE_ADJ && !AA) bits |= (S.q_nc[qs + b] & 0xf) << 4;
          if (c == NC - 1) P.act[b] = bits;
        }
        __syncthreads();
        if constexpr (AA && MODE != MODE_ADJ) {
          // double integration (causal, then anticausal) of each window,
          // unless the traces come integrated
          if (!a.int_a) {
            for (int b = wave; b < nb; b += kWaves)
              integrate_window<SA, SC>(Wa + SA * S.b_off[b], Wc + SC * S.b_off[b], S.q_len[qs + b], lane);
            __syncthreads();
          }
        }

        // -- contributions of the traces that reach this wave's samples
        auto trace_body = [&](const int b, const int base, const float2 (&h2v)[NC], const float (&hmv)[NC],
                              auto full_tag, auto narrow_tag, auto seg_tag) {
          constexpr bool SEG = decltype(seg_tag)::value;
          // kind 1 (FULL): the trace's times stay inside the record in this
          // wave (no clamp); in the forward it also passes every mask.  Kind 2
          // (forward): it passes every mask but may clamp.  Kind 3 (adjoint):
          // inside the record and passes every mask; kind 4 (adjoint): passes
          // every mask but may clamp.  Kind 0: general.
          constexpr int KIND = decltype(full_tag)::value;
          constexpr bool FULL = KIND == 1 || KIND == 3;
          constexpr bool UNMASKED = (KIND != 0 && MODE == MODE_FWD) || KIND == 3 || KIND == 4;
          constexpr bool NARROW = decltype(narrow_tag)::value;
          // adjoint, clamping traces: the contributions read past the record
          // end all land on samples n-2 and n-1; they are summed in a register
          // and scattered once per trace (same-address LDS atomics of many
          // lanes serialise)
          [[maybe_unused]] float cacc = 0.f;
          // window of the trace, indexed by sample
          [[maybe_unused]] float* const Fb = Wf + base;
          [[maybe_unused]] const float2* const Pb = Wp + base;
          [[maybe_unused]] double* const Ab = Wa + SA * base;
          [[maybe_unused]] float* const Cb = Wc + SC * base;
          [[maybe_unused]] unsigned long long* const Lb = reinterpret_cast<unsigned long long*>(Wl + base);
          // antialias: the next column's slopes are read one column ahead
          [[maybe_unused]] float4 gnext = AA ? P.g[b][0] : make_float4(0.f, 0.f, 0.f, 0.f);
#pragma unroll
          for (int c = 0; c < NC; ++c) {
            const float2 h2 = h2v[c];
            const float hm = hmv[c];
            bool ok = UNMASKED || hm < hl2[c];
            // FP32 form of the same predicate for the branch-free forward:
            // sat(2^64 (hlim^2 - hm)) is 1 exactly when hm < hlim^2 (the FMA
            // rounds the exact, power-of-two scaled difference once, so its
            // sign is exact; any positive difference of these magnitudes,
            // hlim^2 >= (1e-6 tan(theta))^2, scales to >= 1).
            [[maybe_unused]] float mk = 1.f;
            if constexpr (MODE != MODE_GRAD && !GEN && !UNMASKED)
              mk = __builtin_amdgcn_fmed3f(fmaf(hm, -0x1p64f, hl2[c] * 0x1p64f), 0.f, 1.f);
            const float xs = fmaf(h2.x, qd[c], tqs), xr = fmaf(h2.y, qd[c], tqs);
            const float rs = __builtin_amdgcn_rsqf(xs), rr = __builtin_amdgcn_rsqf(xr);
            float u_raw;
            if constexpr (VTI) {
              // anelliptic leg, p = h^2 qd (samples^2):
              //   t^2 = tqs + p - 2 eta p^2 / (tqs + (1 + 2 eta) p) = N / D,
              //   N = p^2 + 2 (1 + eta) tqs p + tqs^2,  D = tqs + (1 + 2 eta) p,
              // both sums of positive terms for 1 + 2 eta > 0 (no cancellation,
              // N >= tqs^2 > 0); t = N rsq(N D): one transcendental per leg
              const float ps = h2.x * qd[c], pr = h2.y * qd[c];
              const float ns = fmaf(ps, ps + c1[c], tqs2), nr = fmaf(pr, pr + c1[c], tqs2);
              const float ds = fmaf(k1[c], ps, tqs), dr = fmaf(k1[c], pr, tqs);
              u_raw = fmaf(ns, __builtin_amdgcn_rsqf(ns * ds), fmaf(nr, __builtin_amdgcn_rsqf(nr * dr), -a.t0_over_dt));
            } else {
              u_raw = fmaf(xs, rs, fmaf(xr, rr, -a.t0_over_dt));
            }
            if (zero_out) ok = ok && (u_raw >= 0.f) && (u_raw <= a.u_last);
            const float u = FULL ? u_raw : __builtin_amdgcn_fmed3f(u_raw, 0.f, a.umax);
            const float prs = rs * rr;
            float wd = prs * prs;
            if constexpr (GEN) {
              if (wmode != PSTM_WEIGHT_FULL) wd = wmode == PSTM_WEIGHT_NONE ? 1.f : prs;
              if (use_taper)  // z = v tau / 2 = sqrt(tqs / qd)
                wd *= aperture_taper(h2.x, h2.y, __builtin_amdgcn_sqrtf(tqs * __builtin_amdgcn_rcpf(qd[c])), a);
            } else if constexpr (WM == PSTM_WEIGHT_SPREADING) {
              wd = prs;
            }
            [[maybe_unused]] float L = 1.f, iL2 = 1.f;
            if constexpr (AA) {
              const float4 g = gnext;
              if (c + 1 < NC) gnext = P.g[b][c + 1];
              const float ga = fmaf(g.x, rs, g.z * rr), gb = fmaf(g.y, rs, g.w * rr);
              L = __builtin_amdgcn_fmed3f(qd[c] * (fabsf(ga) + fabsf(gb)), 1.f, a.aa_max);
              iL2 = __builtin_amdgcn_rcpf(L * L);
            }
            // SEG: only sample indices of this segment count; the other
            // lanes read at sg_rd, whose taps lie inside the staged segment
            [[maybe_unused]] float us = u;
            if constexpr (SEG) {
              static_assert(KIND == 0, "segments take the general loop");
              const int it = (int)u;
              const bool ins = it >= sg_lo && it < sg_hi;
              ok = ok && ins;
              mk = ins ? mk : 0.f;
              us = ins ? u : sg_rd;
            }
            const int i0 = (int)us;  // us >= 0: truncation is floor
            const float f = __builtin_amdgcn_fractf(us);
            if constexpr (MODE == MODE_FWD || MODE == MODE_GRAD) {
              float amp;
              [[maybe_unused]] float a0 = 0.f, a1 = 0.f;
              if constexpr (AA) {
                // triangle taps: second difference of the double integral A,
                // A(i + f) = A(i) - f C(i) with C the single integral
                const float um = us - L, up = us + L;
                const float flm = floorf(um);
                const int im = (int)flm, ip = (int)up;
                const float fm = um - flm, fp = __builtin_amdgcn_fractf(up);
                const double a0 = Ab[SA * i0], am = Ab[SA * im], ap = Ab[SA * ip];
                const float c0 = Cb[SC * i0], cm = Cb[SC * im], cp = Cb[SC * ip];
                // the six reads are issued together: one LDS round trip per column
                // (left to itself the scheduler spaces them over three)
                __builtin_amdgcn_sched_barrier(0);
                const double y2 = fma(2.0, a0, -am) - ap;
                // the FP32 corrections -2 f C(i0) + fm C(im) + fp C(ip) chained onto it
                amp = fmaf(-2.f * f, c0, fmaf(fm, cm, fmaf(fp, cp, (float)y2))) * iL2;
              } else if constexpr (PAIR) {
                const float2 e = Pb[i0];
                amp = fmaf(f, e.y, e.x);
              } else {
                a0 = Fb[i0];
                a1 = Fb[i0 + 1];
                amp = fmaf(f, a1 - a0, a0);
              }
              if constexpr (MODE == MODE_FWD && UNMASKED) {
                accb[c] = fmaf(wd, amp, accb[c]);
              } else if constexpr (MODE == MODE_FWD && !GEN) {
                accb[c] = fmaf(wd * mk, amp, accb[c]);
                fcnt[c] += mk;
              } else if constexpr (MODE == MODE_FWD) {
                accb[c] = fmaf(ok ? wd : 0.f, amp, accb[c]);
                fcnt[c] += ok ? 1.f : 0.f;
              } else {
                // d/dv of cw * wd * amp with masks held (isotropic, no antialias)
                const float Q = qd[c] * fmaf(h2.x, rs * rs, h2.y * (rr * rr));  // q (h_s^2/t_s^2 + h_r^2/t_r^2)
                float dwf;  // (d/dv of the weight) / weight
                if (wmode == PSTM_WEIGHT_FULL) dwf = -2.f * sv[c] * (1.f - Q);
                else if (wmode == PSTM_WEIGHT_SPREADING) dwf = -2.f * sv[c] * (1.f - 0.5f * Q);
                else if (wmode == PSTM_WEIGHT_OBLIQUITY) dwf = sv[c] * Q;
                else dwf = 0.f;
                const bool clamped = !(u_raw >= 0.f && u_raw <= a.umax);
                const float dudv = -sv[c] * qd[c] * fmaf(h2.x, rs, h2.y * rr);  // d(t/dt)/dv
                const float slope = clamped ? 0.f : (a1 - a0);
                const float d = wd * fmaf(slope, dudv, dwf * amp);
                accb[c] += ok ? d : 0.f;
              }
            } else {
              // adjoint: scatter w * m into the window (transpose of the taps)
              // branch-free masks for the common form (same FP32 predicate as
              // the forward's); the general form selects
              float gsc = GEN ? (ok ? wd * m_in[c] : 0.f) : wd * m_in[c] * mk;
              if constexpr (!AA && !FULL) {
                // 1 exactly when u_raw > umax (sign of the exact scaled difference)
                const float cl = __builtin_amdgcn_fmed3f(fmaf(u_raw, 0x1p64f, -a.umax * 0x1p64f), 0.f, 1.f);
                cacc = fmaf(gsc, cl, cacc);
                gsc = fmaf(-gsc, cl, gsc);
              }
              // lanes without a contribution skip the atomics: their positions
              // may lie outside the staged window, and clamped lanes share one.
              // Kind 3 needs no skip: every lane of the wave passes the masks
              // and reads inside the window (a zero tap adds nothing).
              if (KIND == 3 || gsc != 0.f) {
                if constexpr (SKEW && NARROW) {
                  // gsc < 2^30 in scaled units; each tap rounded toward zero
                  lds_u64* const q = skewed(Lb, i0);
                  const float t1 = gsc * f, t0 = gsc - t1;
                  lds_add(q, (unsigned long long)(long long)(int)t0);
                  lds_add(q + 1, (unsigned long long)(long long)(int)t1);
                } else if constexpr (SKEW) {
                  // |g| < 2^47 (the workgroup's scale): two-instruction rounding
                  lds_u64* const q = skewed(Lb, i0);
                  const double g = (double)gsc, gf = g * (double)f;
                  lds_add(q, to_fixed64_small(g - gf));
                  lds_add(q + 1, to_fixed64_small(gf));
                } else if constexpr (FX) {
                  const double g = (double)gsc, gf = g * (double)f;
                  if constexpr (AA) {
                    const float um = us - L, up = us + L;
                    const int im = (int)floorf(um), ip = (int)up;
                    const double fm = (double)(um - (float)im), fp = (double)(up - (float)ip);
                    const double gc = 2.0 * g * (double)iL2, ge = -g * (double)iL2, gcf = gc * (double)f;
                    atomicAdd(&Lb[i0], to_fixed64_small(gc - gcf));
                    atomicAdd(&Lb[i0 + 1], to_fixed64_small(gcf));
                    atomicAdd(&Lb[im], to_fixed64_small(ge - ge * fm));
                    atomicAdd(&Lb[im + 1], to_fixed64_small(ge * fm));
                    atomicAdd(&Lb[ip], to_fixed64_small(ge - ge * fp));
                    atomicAdd(&Lb[ip + 1], to_fixed64_small(ge * fp));
                  } else {
                    atomicAdd(&Lb[i0], to_fixed64_small(g - gf));
                    atomicAdd(&Lb[i0 + 1], to_fixed64_small(gf));
                  }
                } else if constexpr (AA) {
                  const float um = us - L, up = us + L;
                  const int im = (int)floorf(um), ip = (int)up;
                  const float fm = um - (float)im, fp = up - (float)ip;
                  const double gc = 2.0 * (double)(gsc * iL2), ge = -(double)(gsc * iL2);
                  atomicAdd(&Ab[i0], gc * (1.0 - f));
                  atomicAdd(&Ab[i0 + 1], gc * f);
                  atomicAdd(&Ab[im], ge * (1.0 - fm));
                  atomicAdd(&Ab[im + 1], ge * fm);
                  atomicAdd(&Ab[ip], ge * (1.0 - fp));
                  atomicAdd(&Ab[ip + 1], ge * fp);
                } else {
                  atomicAdd(&Fb[i0], gsc * (1.f - f));
                  atomicAdd(&Fb[i0 + 1], gsc * f);
                }
              }
            }
          }
          if constexpr (MODE == MODE_ADJ && !AA && !FULL && CLAMPSUM) {
            // row sums (16 lanes each), added by four lanes to the trace's clamp sum
            if (__ballot(cacc != 0.f) != 0) {
              float rsum = cacc;
              rsum += dpp_move<kDppRowShr1>(rsum);
              rsum += dpp_move<kDppRowShr2>(rsum);
              rsum += dpp_move<kDppRowShr4>(rsum);
              rsum += dpp_move<kDppRowShr8>(rsum);
              if ((lane & 15) == 15 && rsum != 0.f) atomicAdd(&clampS[b], rsum);
            }
          } else if constexpr (MODE == MODE_ADJ && !AA && !FULL) {
            if (__ballot(cacc != 0.f) != 0) {
              const float tot = wave_scan(cacc, lane);  // lane 63: the wave's sum
              if (lane == 63 && tot != 0.f) {
                const int ic = (int)a.umax;  // clamped position: sample n-2, fraction 0.999
                const float fc = __builtin_amdgcn_fractf(a.umax);
                const float t1 = tot * fc, t0 = tot - t1;
                if constexpr (FX) {
                  const int pos = SKEW ? ic + (ic >> 5) : ic;  // taps already in scaled units
                  atomicAdd(&Lb[pos], to_fixed64((double)t0));
                  atomicAdd(&Lb[pos + 1], to_fixed64((double)t1));
                } else {
                  atomicAdd(&Fb[ic], t0);
                  atomicAdd(&Fb[ic + 1], t1);
                }
              }
            }
          }
        };
        // Walk the set bits of a trace mask; the next trace's distances are
        // read from LDS before the current trace is summed, so their latency
        // overlaps the arithmetic.
        auto walk = [&](uint64_t m, auto full_tag, auto narrow_tag, auto seg_tag) {
          if constexpr (!PREFETCH) {
            for (; m != 0; m &= m - 1) {
              const int b = uniform_i(__builtin_ctzll(m));
              float2 h2c[NC];
              float hmc[NC];
#pragma unroll
              for (int c = 0; c < NC; ++c) {
                h2c[c] = P.h2[b][c];
                hmc[c] = P.hm[b][c];
              }
              trace_body(b, uniform_i(S.b_base[b]), h2c, hmc, full_tag, narrow_tag, seg_tag);
            }
            return;
          }
          if (m == 0) return;
          int b = uniform_i(__builtin_ctzll(m));
          float2 h2c[NC];
          float hmc[NC];
#pragma unroll
          for (int c = 0; c < NC; ++c) {
            h2c[c] = P.h2[b][c];
            hmc[c] = P.hm[b][c];
          }
          int basec = uniform_i(S.b_base[b]);
          while (true) {
            m &= m - 1;
            const bool more = m != 0;
            const int bn = more ? uniform_i(__builtin_ctzll(m)) : b;
            float2 h2n[NC];
            float hmn[NC];
#pragma unroll
            for (int c = 0; c < NC; ++c) {
              h2n[c] = P.h2[bn][c];
              hmn[c] = P.hm[bn][c];
            }
            const int basen = uniform_i(S.b_base[bn]);
            trace_body(b, basec, h2c, hmc, full_tag, narrow_tag, seg_tag);
            if (!more) break;
            b = bn;
            basec = basen;
#pragma unroll
            for (int c = 0; c < NC; ++c) {
              h2c[c] = h2n[c];
              hmc[c] = hmn[c];
            }
          }
        };
        const uint64_t todo = __ballot(lane < nb && ((P.act[lane] >> wave) & 1));
        uint64_t tfull = 0;
        if (CAN_SEG && seg) {
          if constexpr (CAN_SEG) {
            if (!SKEW || narrow) walk(todo, std::integral_constant<int, 0>{}, std::true_type{}, std::true_type{});
            else walk(todo, std::integral_constant<int, 0>{}, std::false_type{}, std::true_type{});
          }
        } else if constexpr (FULLPATH) {
          tfull = __ballot(lane < nb && ((P.act[lane] >> (4 + wave)) & 1));
          const uint64_t tunm = __ballot(lane < nb && ((P.act[lane] >> (8 + wave)) & 1));
          nfull += (int)__popcll(tfull) + (int)__popcll(tunm);
          walk(tfull, std::integral_constant<int, 1>{}, std::true_type{}, std::false_type{});
          walk(tunm, std::integral_constant<int, 2>{}, std::true_type{}, std::false_type{});
          tfull |= tunm;
        }
        if (CAN_SEG && seg) {
        } else if constexpr (MODE == MODE_ADJ && !AA) {
          // traces inside the record in this wave that pass every mask (ADJ_UNM),
          // the other traces inside the record, then those that clamp
          tfull = __ballot(lane < nb && ((P.act[lane] >> (4 + wave)) & 1));
          const uint64_t tunm = ADJ_UNM ? __ballot(lane < nb && ((P.act[lane] >> (8 + wave)) & 1)) : 0;
          const uint64_t tunc = ADJ_UNM ? __ballot(lane < nb && ((P.act[lane] >> (12 + wave)) & 1)) : 0;
          if (!SKEW || narrow) {
            if constexpr (ADJ_UNM) {
              walk(tunm, std::integral_constant<int, 3>{}, std::true_type{}, std::false_type{});
              walk(tunc, std::integral_constant<int, 4>{}, std::true_type{}, std::false_type{});
            }
            walk(tfull & ~tunm, std::integral_constant<int, 1>{}, std::true_type{}, std::false_type{});
            walk(todo & ~tfull & ~tunc, std::integral_constant<int, 0>{}, std::true_type{}, std::false_type{});
          } else {
            if constexpr (ADJ_UNM) {
              walk(tunm, std::integral_constant<int, 3>{}, std::false_type{}, std::false_type{});
              walk(tunc, std::integral_constant<int, 4>{}, std::false_type{}, std::false_type{});
            }
            walk(tfull & ~tunm, std::integral_constant<int, 1>{}, std::false_type{}, std::false_type{});
            walk(todo & ~tfull & ~tunc, std::integral_constant<int, 0>{}, std::false_type{}, std::false_type{});
          }
        } else {
          walk(todo & ~tfull, std::integral_constant<int, 0>{}, std::true_type{}, std::false_type{});
        }
        if constexpr (MODE != MODE_ADJ) {
          nsum += __popcll(todo);
          if (nsum >= kFlushTraces) {
            join_partials();
            nsum = 0;
          }
        }
        __syncthreads();
        if constexpr (MODE == MODE_ADJ) {
          if constexpr (AA) {
            // transpose of the double integration: the same symmetric operator
            // (fixed-point windows are first converted, exactly up to 2^-53)
            for (int b = wave; b < nb; b += kWaves) {
              if constexpr (FX) {
                const int off = S.b_off[b], len = S.q_len[qs + b];
                for (int s = lane; s < len; s += 64) Wa[off + s] = from_fixed64(Wl[off + s]);
              }
              integrate_window(Wa + S.b_off[b], nullptr, S.q_len[qs + b], lane);
            }
            __syncthreads();
          }
          // flush the windows into the output traces
          for (int b = wave; b < nb; b += kWaves) {
            const int e = qs + b;
            const int w0 = S.q_w0[e], len = S.q_len[e], off = S.b_off[b];
            const int id = a.ids[S.q_j[e]];
            float* row = a.traces_out + (int64_t)id * a.n_samples;
            // transpose of the forward's per-trace offset taper
            const float ow = use_taper ? offset_taper(a.offs[S.q_j[e]], a) : 1.f;
            [[maybe_unused]] const long long* const Lw = Wl + S.b_base[b];
            // contributions that read past the record end: sample n-2 takes
            // (1 - fc) of their sum, n-1 takes fc (fc = the fraction of n - 1.001)
            [[maybe_unused]] float cl_lo = 0.f, cl_hi = 0.f;
            if constexpr (CLAMPSUM) {
              const float fc = __builtin_amdgcn_fractf(a.umax);
              const float cs = clampS[b] * (FX ? (float)fx_inv : 1.f);
              cl_hi = cs * fc;
              cl_lo = cs - cl_hi;
            }
            for (int s = lane; s < len; s += 64) {
              const int dst = w0 + s;
              float v;
              if constexpr (SKEW) {
                const int pos = dst + (dst >> 5);
                const long long t = Lw[pos] + ((dst & 31) == 0 ? Lw[pos - 1] : 0ll);
                v = (float)(from_fixed64(t) * fx_inv);
              } else if constexpr (FX && AA) v = (float)(Wa[off + s] * fx_inv);
              else if constexpr (FX) v = (float)(from_fixed64(Wl[off + s]) * fx_inv);
              else v = AA ? (float)Wa[off + s] : Wf[off + s];
              if constexpr (CLAMPSUM) v += dst == a.n_samples - 2 ? cl_lo : (dst == a.n_samples - 1 ? cl_hi : 0.f);
              if (dst >= 0 && dst < a.n_samples && v != 0.f) atomicAdd(&row[dst], v * ow);
            }
          }
          __syncthreads();
        }
      };
      // A window longer than the LDS share (only possible in the forms a
      // dispatch gives traces longer than every other form's share: CAN_SEG)
      // is processed in segments of at most seg_cap staged elements, each
      // assigned the sample indices [lo, hi) and staged with the taps' reach
      // on either side (the next sample; with antialias the triangle's
      // half-width and its neighbour), clipped to the window.  One call site
      // of run_batch serves both kinds of sub-batch.
      constexpr int seg_cap = SKEW ? ((CAPE - 2) * 32) / 33 : CAPE;
      // with antialias (pad <= 258) every segment holds at least pad indices
      // beyond the taps' reach on both sides
      static_assert(!CAN_SEG || !AA || seg_cap >= 4 * 259 + 64, "LDS share too small for antialias segments");
      static_assert(!CAN_SEG || seg_cap >= 64, "LDS share too small for segments");
      bool segmenting = false;
      int seg_w0 = 0, seg_lo = 0, seg_step = 0, seg_end = 0;
      const int pad_lo = AA ? a.aa_pad + 1 : 0, pad_hi = AA ? a.aa_pad + 2 : 1;
      while (true) {
        int nb = 0;
        if (!segmenting) {
          const int rem = qn - qs;
          if (rem <= 0 || (!final_flush && rem < MB)) break;
          if (wave == 0) {
            const int i = qs + lane;
            const bool in = lane < MB && i < qn;
            const int cost = in ? (SKEW ? skew_len(S.q_len[i]) : S.q_len[i]) : 0;  // window elements
            const int incl = wave_scan(cost, lane);
            const bool fit = in && incl <= CAPE;
            const uint64_t m = __ballot(fit);
            if (fit) {
              const int off = incl - cost;
              const int w0 = S.q_w0[i];
              S.b_off[lane] = off;
              // SKEW: sample i at base + i + (i >> 5), one spare slot before the window (w0 >= 0)
              S.b_base[lane] = SKEW ? off + 1 - (w0 + (w0 >> 5)) : off - w0;
            }
            if (lane == 0) S.nb = __popcll(m);
          }
          __syncthreads();
          nb = S.nb;
          if (nb == 0) {
            if constexpr (!CAN_SEG) {
              qs += 1;  // not reached: the dispatch gives these forms windows that fit
              continue;
            }
            // near-equal segments (none short) of `seg_step` indices
            const int w0 = S.q_w0[qs], len = S.q_len[qs];
            const int cap_idx = seg_cap - pad_lo - pad_hi;
            const int n_seg = (len + cap_idx - 1) / cap_idx;
            seg_step = (len + n_seg - 1) / n_seg;
            seg_w0 = w0;
            seg_lo = w0;
            seg_end = w0 + len;
            segmenting = true;
          }
        }
        int lo = 0, hi = 0;
        float rd = 0.f;
        if (CAN_SEG && segmenting) {
          lo = seg_lo;
          hi = min(lo + seg_step, seg_end);
          const int s0 = max(lo - pad_lo, seg_w0), s1 = min(hi + pad_hi, seg_end);
          // read position of the lanes outside the segment: its taps (the
          // next sample; with antialias +-(aa_pad + 1)) stay in [s0, s1)
          rd = (float)(AA ? s0 + pad_lo : lo);
          __syncthreads();  // every thread has read the entry before it is rewritten
          if (tid == 0) {
            S.q_w0[qs] = s0;
            S.q_len[qs] = s1 - s0;
            S.b_off[0] = 0;
            S.b_base[0] = SKEW ? 1 - (s0 + (s0 >> 5)) : -s0;
          }
          __syncthreads();
          nb = 1;
        }
        run_batch(qs, nb, CAN_SEG && segmenting, lo, hi, rd);
        if (CAN_SEG && segmenting) {
          seg_lo = hi;
          if (seg_lo >= seg_end) {
            segmenting = false;
            qs += 1;
          }
        } else {
          qs += nb;
        }
      }
      // move the unprocessed tail (< MB entries) to the front
      const int rem = qn - qs;
      int j = 0, w0 = 0, len = 0, nc = 0;
      if (tid < rem) {
        j = S.q_j[qs + tid];
        w0 = S.q_w0[qs + tid];
        len = S.q_len[qs + tid];
        nc = S.q_nc[qs + tid];
      }
      __syncthreads();
      if (tid < rem) {
        S.q_j[tid] = j;
        S.q_w0[tid] = w0;
        S.q_len[tid] = len;
        S.q_nc[tid] = (uint8_t)nc;
      }
      if (tid == 0) S.queue_n = rem > 0 ? rem : 0;
      __syncthreads();
    };

    const int n_cells = a.ncx * a.ncy;
    for (int band = a.band0; band < a.band1; ++band)
    for (int cy = cy0; cy <= cy1; ++cy) {
      const int r0 = a.cell_start[band * n_cells + cy * a.ncx + cx0];
      const int r1 = a.cell_start[band * n_cells + cy * a.ncx + cx1 + 1];
      for (int j0 = r0; j0 < r1; j0 += kThreads) {
        const int j = j0 + tid;
        bool pass = false;
        int nclamp = 0;  // bit w: wave w reads inside the record only
        int w0 = 0, len = 0;
        if (j < r1) {
          const float4 g = a.geom[j];
          // distances from source / receiver to the tile rectangle
          const float sdx = fmaxf(fmaxf(X0 - g.x, g.x - X1), 0.f), sdy = fmaxf(fmaxf(Y0 - g.y, g.y - Y1), 0.f);
          const float rdx = fmaxf(fmaxf(X0 - g.z, g.z - X1), 0.f), rdy = fmaxf(fmaxf(Y0 - g.w, g.w - Y1), 0.f);
          const float hs_min = fmaf(sdx, sdx, sdy * sdy), hr_min = fmaf(rdx, rdx, rdy * rdy);
          if (hs_min < hblk && hr_min < hblk) {
            const float sfx = fmaxf(fabsf(g.x - X0), fabsf(g.x - X1)), sfy = fmaxf(fabsf(g.y - Y0), fabsf(g.y - Y1));
            const float rfx = fmaxf(fabsf(g.z - X0), fabsf(g.z - X1)), rfy = fmaxf(fabsf(g.w - Y0), fabsf(g.w - Y1));
            const float hs_max = fmaf(sfx, sfx, sfy * sfy), hr_max = fmaf(rfx, rfx, rfy * rfy);
            // anelliptic times: lower bound with the largest eta >= 0, upper
            // bound with the smallest eta <= 0 (isotropic when eta >= 0)
            const float ecorr = 1.0f / (1.0f + 2.0f * emax), ecorr_hi = 1.0f / (1.0f + 2.0f * emin);
            float tlo = 3.0e38f, thi = -1.f;
#pragma unroll
            for (int w = 0; w < kWaves; ++w) {
              const float hw = S.hwave[w];
              if (hs_min < hw && hr_min < hw) {
                const float ta = fmaxf(S.tqa[w], 1e-12f), tbq = S.tqb[w] + 1e-12f;
                // hardware square roots (1 ulp, far inside the window margin)
                const float lo = __builtin_amdgcn_sqrtf(ta + hs_min * S.qmin[w] * ecorr) +
                                 __builtin_amdgcn_sqrtf(ta + hr_min * S.qmin[w] * ecorr);
                const float hi = __builtin_amdgcn_sqrtf(tbq + hs_max * S.qmax[w] * ecorr_hi) +
                                 __builtin_amdgcn_sqrtf(tbq + hr_max * S.qmax[w] * ecorr_hi);
                tlo = fminf(tlo, lo);
                thi = fmaxf(thi, hi);
                // wave w never reads past either end of the record (the staged
                // window already trusts these bounds to kWindowMargin samples)
                const bool in_rec = fmaf(hi, a.inv_dt, -a.t0_over_dt) + (float)kWindowMargin < a.umax &&
                                    (a.t0_over_dt == 0.f || fmaf(lo, a.inv_dt, -a.t0_over_dt) > (float)kWindowMargin);
                nclamp |= in_rec ? (1 << w) : 0;
              }
            }
            if (thi >= 0.f) {
              const float ul = __builtin_amdgcn_fmed3f(fmaf(tlo, a.inv_dt, -a.t0_over_dt), 0.f, a.umax);
              const float uh = __builtin_amdgcn_fmed3f(fmaf(thi, a.inv_dt, -a.t0_over_dt), 0.f, a.umax);
              int lo_i = (int)ul - pad - kWindowMargin;
              int hi_i = (int)uh + 1 + pad + kWindowMargin;
              lo_i = max(lo_i, AA ? wlo : 0);
              hi_i = min(hi_i, AA ? whi : a.n_samples - 1);
              w0 = lo_i;
              len = hi_i - lo_i + 1;
              pass = true;
            }
          }
        }
        // ordered compaction into the queue
        const uint64_t m = __ballot(pass);
        const int below = __popcll(m & ((1ull << lane) - 1ull));
        if (lane == 0) S.scan[wave] = __popcll(m);
        __syncthreads();
        int pre = 0, tot = 0;
#pragma unroll
        for (int w = 0; w < kWaves; ++w) {
          const int cnt = S.scan[w];
          pre += w < wave ? cnt : 0;
          tot += cnt;
        }
        const int qn = S.queue_n;
        if (pass) {
          const int slot = qn + pre + below;
          S.q_j[slot] = j;
          S.q_w0[slot] = w0;
          S.q_len[slot] = len;
          S.q_nc[slot] = (uint8_t)nclamp;
        }
        __syncthreads();
        if (tid == 0) S.queue_n = qn + tot;
        __syncthreads();
        if (qn + tot >= MB) process_queue(false);
      }
    }
    process_queue(true);
    if constexpr (MODE != MODE_ADJ) join_partials();
  }

  // ---- write-out
  if (!kvalid) return;
#pragma unroll
  for (int c = 0; c < NC; ++c) {
    if (!((cvalid >> c) & 1u)) continue;
    const int il = tx * TX + c / TY, xl = ty * TY + c % TY;
    const int64_t idx = ((int64_t)il * a.n_xl + xl) * a.n_tau + k;
    if constexpr (MODE == MODE_FWD) {
      if (a.comp) {
        float s = a.image[idx], cc = a.comp[idx];
        two_sum_add(s, cc, acc[c]);
        a.image[idx] = s;
        a.comp[idx] = cc + cmp[c];
      } else {
        a.image[idx] += acc[c] + cmp[c];
      }
      if (a.fold) a.fold[idx] += fcnt[c] + (float)nfull;
    } else if constexpr (MODE == MODE_GRAD) {
      a.vel_grad[idx] += (acc[c] + cmp[c]) * a.image_in[idx];
    }
  }
}

// ---------------------------------------------------------------------------
// Host-side launch of one instantiation.

inline int64_t tile_blocks(const TileArgs& a) {
  const int n_tiles = a.n_tiles_x * a.n_tiles_y;
  const int chunk = (n_tiles + 7) / 8;
  return (int64_t)8 * chunk * a.n_tblocks;
}

template <int MODE, int TX, int TY, int DT, bool AA, bool VTI, bool GEN, int OCC = default_occ(AA),
          bool FIXED = false, bool V1 = false, int WM = PSTM_WEIGHT_FULL, bool SEG = false>
int launch_one(const TileArgs& a, hipStream_t s) {
  static_assert(!GEN || WM == PSTM_WEIGHT_FULL, "the general forms read the weight mode");
  const int64_t blocks = tile_blocks(a);
  if (blocks <= 0) return PSTM_OK;
  if (blocks > 0x7fffffff) return set_error(PSTM_ERR_INVALID, "too many tiles (%lld)", (long long)blocks);
  // one launch per offset band (stream-ordered, accumulating)
  TileArgs b = a;
  for (int band = a.band0; band < a.band1; ++band) {
    b.band0 = band;
    b.band1 = band + 1;
    hipLaunchKernelGGL((k_tile<MODE, TX, TY, DT, AA, VTI, GEN, OCC, FIXED, V1, WM, SEG>), dim3((unsigned)blocks),
                       dim3(kThreads), 0, s, b);
  }
  return PSTM_OK;
}

// Whether a call can use the compile-time fast path (GEN = false).
inline bool fast_path(const pstm_params_t* p) {
  return p->weight_mode == PSTM_WEIGHT_FULL && !(p->flags & (PSTM_FLAG_TAPER | PSTM_FLAG_ZERO_OUTSIDE));
}
// The plain forward also has fast forms for the SPREADING weight.
inline bool fast_path_spreading(const pstm_params_t* p) {
  return p->weight_mode == PSTM_WEIGHT_SPREADING && !(p->flags & (PSTM_FLAG_TAPER | PSTM_FLAG_ZERO_OUTSIDE));
}

// Whether a call's whole trace window (every sample plus the antialias
// padding) may exceed the LDS share of the lowest-occupancy forms (3
// workgroups per CU plain, 2 with antialias; 8-column tiles): such calls run
// the segmenting general forms.
inline bool long_traces(const pstm_params_t* p, const TileArgs& a) {
  const bool aa = (p->flags & PSTM_FLAG_AA) != 0;
  return whole_window(p, a.aa_pad) > min_cap_elems(aa);
}

// Per-dtype forward dispatch (defined in pstm_fwd_<dtype>.hip).
template <int DT>
int launch_forward_dt(const pstm_params_t* p, const PlanLayout& L, const TileArgs& a, hipStream_t s);

}  // namespace pstm

namespace pstm {

// Forward dispatch over tile shape, antialias, VTI and fast/general path for
// one trace dtype.  The line shapes (2D images) use the general path only.
template <int DT>
int launch_forward_impl(const pstm_params_t* p, const PlanLayout& L, const TileArgs& a, hipStream_t s) {
  const bool aa = (p->flags & PSTM_FLAG_AA) != 0, vti = (p->flags & PSTM_FLAG_VTI) != 0;
  const bool gen = !fast_path(p);
  if (long_traces(p, a)) {
    // windows may exceed every form's LDS share: the general form at the
    // lowest occupancy, staging such windows in segments
#define PSTM_FWDS(TX, TY)                                                                                     \
  (aa ? (vti ? launch_one<MODE_FWD, TX, TY, DT, true, true, true, 2, false, false, PSTM_WEIGHT_FULL, true>(a, s)    \
             : launch_one<MODE_FWD, TX, TY, DT, true, false, true, 2, false, false, PSTM_WEIGHT_FULL, true>(a, s))   \
      : (vti ? launch_one<MODE_FWD, TX, TY, DT, false, true, true, 3, false, false, PSTM_WEIGHT_FULL, true>(a, s)   \
             : launch_one<MODE_FWD, TX, TY, DT, false, false, true, 3, false, false, PSTM_WEIGHT_FULL, true>(a, s)))
    if (a.tile_x == 4 && a.tile_y == 2) return PSTM_FWDS(4, 2);
    if (a.tile_x == 8 && a.tile_y == 1) return PSTM_FWDS(8, 1);
    if (a.tile_x == 1 && a.tile_y == 8) return PSTM_FWDS(1, 8);
#undef PSTM_FWDS
    return set_error(PSTM_ERR_INVALID, "unsupported tile shape %dx%d", a.tile_x, a.tile_y);
  }
#define PSTM_FWD(TX, TY, AAV, VTIV, GENV) launch_one<MODE_FWD, TX, TY, DT, AAV, VTIV, GENV>(a, s)
#define PSTM_FWD4(TX, TY, GENV)                         \
  (aa ? (vti ? PSTM_FWD(TX, TY, true, true, GENV)       \
             : PSTM_FWD(TX, TY, true, false, GENV))     \
      : (vti ? PSTM_FWD(TX, TY, false, true, GENV)      \
             : PSTM_FWD(TX, TY, false, false, GENV)))
  if (a.tile_x == 4 && a.tile_y == 2 && aa && !gen &&
      whole_window(p, a.aa_pad) <= aa_fast_cap_elems<8, 4>()) {
    // short traces: antialias fast path at 4 workgroups per CU
    return vti ? launch_one<MODE_FWD, 4, 2, DT, true, true, false, 4>(a, s)
               : launch_one<MODE_FWD, 4, 2, DT, true, false, false, 4>(a, s);
  }
  if (a.tile_x == 4 && a.tile_y == 2 && aa && !gen &&
      whole_window(p, a.aa_pad) <= aa_fast_cap_elems<8>()) {
    // antialias fast path at 3 workgroups per CU
    return vti ? launch_one<MODE_FWD, 4, 2, DT, true, true, false, 3>(a, s)
               : launch_one<MODE_FWD, 4, 2, DT, true, false, false, 3>(a, s);
  }
  // Plain forward at high occupancy (measured on the 1M-trace workload: 1D
  // velocity at 7 workgroups per CU 1930 Gcontrib/s, 6: 1885, 5: 1849, 4 with
  // prefetch 1814, 3 with prefetch 1624, 8: 1870; 3D velocity on 8 of its 25
  // classes, sample-and-difference windows: 6 per CU 1828, 5: 1823, 4: 1707).
  // A sub-batch must hold a whole trace's window, so the smaller LDS shares
  // serve traces that fit.
  const int64_t win = whole_window(p, 0);
  if (a.tile_x == 4 && a.tile_y == 2 && !gen && !aa && !vti) {
    if (p->vel_mode == PSTM_VEL_1D && win <= pair_cap_elems<7>())
      return launch_one<MODE_FWD, 4, 2, DT, false, false, false, 7, false, true>(a, s);
    if (win <= pair_cap_elems<6>()) return launch_one<MODE_FWD, 4, 2, DT, false, false, false, 6>(a, s);
    if (win <= pair_cap_elems<4>()) return launch_one<MODE_FWD, 4, 2, DT, false, false, false, 4>(a, s);
  }
  // VTI without antialias: 4 per CU where the window fits (864 Gcontrib/s
  // against 820 at 3 on 8 classes of the 1M-trace workload; at 6 per CU the
  // inner loops spill)
  if (a.tile_x == 4 && a.tile_y == 2 && !gen && !aa && vti && win <= sample_cap_elems<8, false, 4>())
    return launch_one<MODE_FWD, 4, 2, DT, false, true, false, 4>(a, s);
  if (a.tile_x == 4 && a.tile_y == 2 && fast_path_spreading(p) && !aa && !vti) {
    constexpr int SP = PSTM_WEIGHT_SPREADING;
    if (p->vel_mode == PSTM_VEL_1D && win <= pair_cap_elems<7>())
      return launch_one<MODE_FWD, 4, 2, DT, false, false, false, 7, false, true, SP>(a, s);
    if (win <= pair_cap_elems<6>()) return launch_one<MODE_FWD, 4, 2, DT, false, false, false, 6, false, false, SP>(a, s);
    if (win <= pair_cap_elems<4>()) return launch_one<MODE_FWD, 4, 2, DT, false, false, false, 4, false, false, SP>(a, s);
  }
  // General forms (taper, zero-outside, NONE or OBLIQUITY weight): 4 per CU
  // where the window fits (127 VGPRs either way; on 8 classes of the 1M-trace
  // workload the cosine taper 590 Gcontrib/s against 529 at 3, OBLIQUITY 1332
  // against 1118)
  if (a.tile_x == 4 && a.tile_y == 2 && gen && !aa && !vti && win <= sample_cap_elems<8, false, 4>())
    return launch_one<MODE_FWD, 4, 2, DT, false, false, true, 4>(a, s);
  if (a.tile_x == 4 && a.tile_y == 2) return gen ? PSTM_FWD4(4, 2, true) : PSTM_FWD4(4, 2, false);
  if (a.tile_x == 8 && a.tile_y == 1) return PSTM_FWD4(8, 1, true);
  if (a.tile_x == 1 && a.tile_y == 8) return PSTM_FWD4(1, 8, true);
#undef PSTM_FWD4
#undef PSTM_FWD
  return set_error(PSTM_ERR_INVALID, "unsupported tile shape %dx%d", a.tile_x, a.tile_y);
}

}  // namespace pstm
