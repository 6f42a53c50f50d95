/*
 * pstm_abi.h -- C ABI of the MI355X Kirchhoff prestack time migration (PSTM)
 * library libpstm_hip.so (gfx950).
 *
 * The reference (SeisProc) runs this path as Python/PyTorch loops with no
 * foreign-function layer.  Each entry point below names the reference
 * function whose work it takes over; the Python host layer
 * (seisproc_amd.processors.kernel_backend) binds these symbols with ctypes.
 *
 * Conventions
 *   - Every pointer argument is a DEVICE pointer (HBM) unless noted, owned by
 *     the caller.  The library allocates no device memory: all persistent
 *     per-job state lives in the caller's plan workspace.
 *   - `stream` is a hipStream_t passed as void* (NULL = the null stream).
 *     Work is enqueued and the call returns without synchronising.
 *   - Return value: PSTM_OK (0) or a negative pstm_status_t.  The message of
 *     the last failure on the calling thread is in pstm_last_error().
 *   - Image, fold, velocity and eta volumes are column-major in time:
 *     index = (il * n_xl + xl) * n_tau + k   ("[il][xl][tau]").
 *   - Traces are trace-major: index = trace * n_samples + sample.
 *   - Times are in seconds, distances in metres, angles in degrees.
 */
#ifndef PSTM_ABI_H
#define PSTM_ABI_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define PSTM_ABI_VERSION 3

typedef enum {
  PSTM_OK = 0,
  PSTM_ERR_INVALID = -1,      /* bad parameter or size */
  PSTM_ERR_UNSUPPORTED = -2,  /* flag combination not available */
  PSTM_ERR_WORKSPACE = -3,    /* workspace too small or not built by pstm_plan_build */
  PSTM_ERR_HIP = -4           /* HIP runtime error (message has the HIP error string) */
} pstm_status_t;

/* Storage type of input traces. */
typedef enum { PSTM_F32 = 0, PSTM_BF16 = 1, PSTM_F16 = 2 } pstm_dtype_t;

/* Behaviour flags (bitwise OR). */
enum {
  PSTM_FLAG_AA = 1u << 0,            /* triangle antialias filter on integrated traces */
  PSTM_FLAG_VTI = 1u << 1,           /* per-leg anelliptic eta traveltime */
  PSTM_FLAG_KAHAN = 1u << 2,         /* compensated FP32 accumulation */
  PSTM_FLAG_TAPER = 1u << 3,         /* cosine edge tapers: distance and angle of both legs, offset window */
  PSTM_FLAG_ZERO_OUTSIDE = 1u << 4,  /* drop contributions outside the record (default: clamp) */
  PSTM_FLAG_TIME_APERTURE = 1u << 5  /* aperture(z) = min(aperture_m, z*tan(max_angle)) */
};

/* Amplitude weight (reference: kirchhoff_kernel.py:691-693 is FULL =
 * cos_s cos_r / (r_s r_r); kirchhoff_migrator.py:378 is SPREADING =
 * 1 / (r_s r_r); OBLIQUITY = cos_s cos_r; NONE = 1;
 * models/migration_config.py:40-45 names the four modes). */
enum { PSTM_WEIGHT_FULL = 0, PSTM_WEIGHT_SPREADING = 1, PSTM_WEIGHT_NONE = 2, PSTM_WEIGHT_OBLIQUITY = 3 };

/* Velocity / eta layout. 1D: one value per output tau, shared by all columns.
 * 3D: one value per image sample, [il][xl][tau]. */
enum { PSTM_VEL_1D = 0, PSTM_VEL_3D = 1 };

/* Job description.  Plain data; identical on host and device. */
typedef struct {
  int64_t n_traces;
  int32_t n_samples;   /* samples per input trace */
  int32_t n_il;        /* output inlines */
  int32_t n_xl;        /* output crosslines */
  int32_t n_tau;       /* output time samples */
  double dt;           /* input sample interval (s) */
  double t0;           /* time of input sample 0 (s) */
  double dtau;         /* output sample interval (s) */
  double tau0;         /* first output time (s) */
  double origin_x;     /* world X of image column (il=0, xl=0) */
  double origin_y;     /* world Y of image column (il=0, xl=0) */
  double il_dx, il_dy; /* world step per inline index */
  double xl_dx, xl_dy; /* world step per crossline index (must be orthogonal to the inline step) */
  double aperture_m;   /* maximum source/receiver-to-image horizontal distance */
  double max_angle_deg;/* maximum angle from vertical, per leg */
  double taper_width;  /* fraction of each limit tapered by PSTM_FLAG_TAPER, in (0, 1) */
  double aa_scale;     /* triangle half-width multiplier (PSTM_FLAG_AA) */
  double aa_max;       /* maximum triangle half-width in samples (PSTM_FLAG_AA) */
  int32_t dtype;       /* pstm_dtype_t of the traces */
  int32_t flags;       /* PSTM_FLAG_* */
  int32_t weight_mode; /* PSTM_WEIGHT_* */
  int32_t vel_mode;    /* PSTM_VEL_* (also the layout of eta) */
  /* Offset window (source-receiver distance), a property of the plan:
   * with max_offset_m > 0 pstm_plan_build keeps only traces with
   * min_offset_m <= offset <= max_offset_m (reference
   * processors/migration/aperture.py:184-224), and PSTM_FLAG_TAPER adds
   * its near and far cosine tapers.  max_offset_m <= 0: no window. */
  double min_offset_m;
  double max_offset_m;
} pstm_params_t;

/* Version of this ABI (PSTM_ABI_VERSION). */
int pstm_abi_version(void);

/* Message of the last failure on this thread ("" if none).  Host pointer,
 * valid until the next call on the same thread. */
const char* pstm_last_error(void);

/* Validate a job description without touching the GPU. */
int pstm_validate(const pstm_params_t* p);

/* Bytes of device workspace a plan for `p` needs. 0 on invalid parameters.
 * Replaces the per-tile trace pre-filter of the reference,
 * processors/migration/kirchhoff_kernel.py:579-590. */
size_t pstm_plan_bytes(const pstm_params_t* p);

/* Build the plan in `ws`: reduces the FP64 world coordinates of every trace
 * to FP32 image-grid-local coordinates (once per trace), drops traces whose
 * midpoint cannot reach the image, and sorts the rest into spatial cells.
 * src_xy, rcv_xy: device, [n_traces][2] FP64 world (x, y).
 * Replaces the FP32 casts and grid set-up of the reference,
 * processors/migration/migration_engine.py:286-301. */
int pstm_plan_build(const pstm_params_t* p, const double* src_xy, const double* rcv_xy,
                    void* ws, size_t ws_bytes, void* stream);

/* Number of traces the plan kept (host value; synchronises `stream`). */
int64_t pstm_plan_kept(const void* ws, void* stream);

/* Convert FP32 traces to BF16/FP16 storage with one FP32 scale per trace
 * (scale = max|trace|, stored = trace/scale, round to nearest even). */
int pstm_quantize_traces(const float* traces, int64_t n_traces, int32_t n_samples,
                         int32_t dtype, void* out, float* scale, void* stream);

/* Forward migration (Kirchhoff summation), accumulating into image/fold.
 *   traces:      [n_traces][n_samples] of p->dtype
 *   trace_scale: [n_traces] FP32 or NULL (all 1)
 *   vel:         v_rms, FP32, layout p->vel_mode
 *   eta:         FP32, layout p->vel_mode, required with PSTM_FLAG_VTI, else ignored;
 *                every value must satisfy 1 + 2 eta > 0, and the isotropic leg times
 *                must stay below 2^21 samples: ((tau_max / 2)^2 + (aperture / v_min)^2)
 *                / dt^2 * ((1 + e) (1 + 2 e))^(1/3) < 2^42, e = max(eta_max, 0) (neither
 *                is checked on the device; the Python layer checks both)
 *   image:       FP32 [il][xl][tau], image += sum of contributions
 *   image_comp:  FP32 [il][xl][tau] running compensation (PSTM_FLAG_KAHAN), or NULL
 *                (then the compensation is folded into image)
 *   fold:        FP32 [il][xl][tau], fold += number of contributions, or NULL
 * Replaces processors/migration/kirchhoff_kernel.py:459-854 (migrate_kirchhoff_full). */
int pstm_forward(const pstm_params_t* p, const void* ws, const void* traces,
                 const float* trace_scale, const float* vel, const float* eta,
                 float* image, float* image_comp, float* fold, void* stream);

/* Number of offset bands a plan for `p` sorts its traces into (the plan
 * orders traces by offset band, then midpoint cell), or a negative status. */
int pstm_plan_bands(const pstm_params_t* p);

/* pstm_forward restricted to the traces of offset bands [band_begin,
 * band_end).  Calls over disjoint band ranges accumulate into the same
 * image, compensation and fold exactly as one pstm_forward call over their
 * union, so a caller can report progress (the reference's progress
 * callback, processors/migration/kirchhoff_kernel.py:760-774) or
 * checkpoint between them. */
int pstm_forward_bands(const pstm_params_t* p, const void* ws, int32_t band_begin, int32_t band_end,
                       const void* traces, const float* trace_scale, const float* vel, const float* eta,
                       float* image, float* image_comp, float* fold, void* stream);

/* Causal and anticausal integrals of every trace, the inputs of
 * pstm_forward_integrated: int_c[t][s] = scale[t] * sum_{j <= s} trace[t][j]
 * (FP32) and int_a[t][s] = sum_{k >= s} C[t][k] (FP64), [n_traces][n_samples]
 * each (12 bytes per sample).  trace_scale may be NULL (all 1).  The triangle
 * antialias filter works on these integrals ("triangle antialias via
 * causal/anticausal integrated traces"); the reference's antialias module
 * (processors/migration/antialias.py) computes filter weights only and is not
 * wired into its migration path, so this has no counterpart there. */
int pstm_integrate_traces(const pstm_params_t* p, const void* traces, const float* trace_scale,
                          double* int_a, float* int_c, void* stream);

/* pstm_forward_bands with the antialias filter (PSTM_FLAG_AA required) on
 * traces integrated once by pstm_integrate_traces, instead of integrating each
 * staged window in LDS for every image tile: same operator (the triangle
 * taps are second differences, unchanged by the linear terms a later
 * integration start adds), less work per tile.  Same bands, accumulation and
 * outputs as pstm_forward_bands (kirchhoff_kernel.py:459-854). */
int pstm_forward_integrated(const pstm_params_t* p, const void* ws, int32_t band_begin, int32_t band_end,
                            const double* int_a, const float* int_c, const float* vel, const float* eta,
                            float* image, float* image_comp, float* fold, void* stream);

/* Adjoint of pstm_forward (Kirchhoff demigration / modelling):
 * traces_out += L^T image, FP32 [n_traces][n_samples].  Same masks, weights,
 * interpolation and antialias taps as the forward operator.
 * No reference counterpart (the reference has no adjoint). */
int pstm_adjoint(const pstm_params_t* p, const void* ws, const float* image,
                 const float* vel, const float* eta, float* traces_out, void* stream);

/* Velocity gradient: vel_grad[il][xl][tau] += image_grad * d(image)/d(v_rms)
 * at every image sample, with the aperture/angle masks held fixed.  Traces as
 * in pstm_forward.  Isotropic traveltime, no antialias (others return
 * PSTM_ERR_UNSUPPORTED).  vel must be PSTM_VEL_3D.
 * No reference counterpart. */
int pstm_vel_grad(const pstm_params_t* p, const void* ws, const void* traces,
                  const float* trace_scale, const float* vel, const float* image_grad,
                  float* vel_grad, void* stream);

/* Size of the static plan workspace of pstm_migrate_static (per GPU). */
#define PSTM_STATIC_SCRATCH_BYTES ((size_t)32 << 20)

/* pstm_plan_build followed by pstm_forward with the plan kept in a static
 * device array of PSTM_STATIC_SCRATCH_BYTES instead of a caller workspace:
 * the call allocates no memory (no hipMalloc, hipFree or allocating library
 * routine), for callers that cannot allocate (code under HIP graph capture,
 * tight serving loops, harnesses that forbid allocation).  No counterpart in
 * the reference.  Arguments as pstm_plan_build and pstm_forward; returns
 * PSTM_ERR_WORKSPACE when the plan of p->n_traces traces needs more than the
 * static workspace (about 56 bytes per trace plus the sort and cell tables:
 * some 570,000 traces for a 400 x 400-column image with a 1500 m aperture;
 * split larger inputs over several calls, which accumulate).  Every call on a GPU shares that GPU's array: calls that
 * overlap on one GPU (other streams or threads) race. */
int pstm_migrate_static(const pstm_params_t* p, const double* src_xy, const double* rcv_xy,
                        const void* traces, const float* trace_scale, const float* vel, const float* eta,
                        float* image, float* image_comp, float* fold, void* stream);

#ifdef __cplusplus
}
#endif

#endif /* PSTM_ABI_H */
