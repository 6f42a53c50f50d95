/*
 * pstm_oracle.c -- FP64 CPU restatement of the reference Kirchhoff PSTM
 * method.  TEST INFRASTRUCTURE ONLY: tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py use it as the checker of the HIP library.  It
 * is never linked into, loaded by, or called from the product path.
 *
 * Reference method followed (paths relative to the SeisProc tree):
 *   processors/migration/kirchhoff_kernel.py:459-854  migrate_kirchhoff_full
 *     - horizontal distances to source/receiver            :611-617
 *     - hard aperture h < A on both legs                    :635
 *     - time-dependent aperture min(A, z tan(theta))        :673-678
 *     - angle mask atan2(h, z + 1e-6) < theta, both legs    :670-671, :683-685
 *     - DSR time (r_s + r_r)/v, r = sqrt(h^2 + z^2)         :661-662, :688
 *     - weight (z/(r_s+e))(z/(r_r+e))/(r_s r_r + e)         :691-693
 *     - clamp to [0, n-1.001], floor/frac lerp              :708-721
 *     - image += sum(a w mask), fold += sum(mask)           :724-730
 *   processors/migration/migration_engine.py:292-306  grid x = x0 + il*d_il,
 *     y = y0 + xl*d_xl, z = v*tau/2
 *   processors/migration/kirchhoff_migrator.py:378    SPREADING weight
 *     1/(sqrt(r_s^2 r_r^2) + 1e-6); OBLIQUITY weight (z/(r_s+e))(z/(r_r+e))
 *     (the cosine factors of :691-693 alone; mode named in
 *     models/migration_config.py:40-45)
 *   processors/migration/traveltime_vti.py:164-222    anelliptic eta time per
 *     leg (v0 = v_rms, delta = 0, epsilon = eta)
 *   processors/migration/aperture.py:147-270          taper composition (distance
 *     and angle tapers of both legs, weights.py:192-236 shape) and the
 *     offset window with near/far tapers (:184-224)
 *
 * Extensions with no reference code (defined here, pinned by the tests):
 *   - v_rms(il, xl, tau): the reference's constant v generalised per image
 *     sample (z = v tau / 2 keeps the reference formulas unchanged);
 *   - triangle antialias: y(u) = (2A(u) - A(u-L) - A(u+L)) / L^2 where A is
 *     the causal-then-anticausal double integral of the zero-extended trace,
 *     linearly interpolated; L = clamp(aa_scale (|dt/dil| + |dt/dxl|) / dt,
 *     1, aa_max) samples.  L = 1 reproduces plain linear interpolation;
 *   - exact adjoint (transpose) and the analytic velocity gradient with the
 *     masks held at `mask_vel`.
 * Parity pins: tests/test_oracle_*.py (known answers restated from the
 * reference tests, brute force, closed forms, dot-product and finite
 * differences).
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#include "../include/pstm_abi.h"

#define EPS_R 1e-6   /* kirchhoff_kernel.py:670-671, :691-693 */

static int g_exhaustive = 0;
void oracle_set_exhaustive(int on) { g_exhaustive = on; }

typedef struct {
  const pstm_params_t* p;
  double tan_th, th_rad;
  double aa_pad;     /* extension of the integrated trace on each side (samples) */
  int pad;
} ctx_t;

static void ctx_init(ctx_t* c, const pstm_params_t* p) {
  c->p = p;
  c->th_rad = p->max_angle_deg * M_PI / 180.0;
  c->tan_th = tan(c->th_rad);
  c->pad = (int)ceil(p->aa_max) + 2;
  c->aa_pad = c->pad;
}

static inline size_t vidx(const pstm_params_t* p, int il, int xl, int k) {
  if (p->vel_mode == PSTM_VEL_1D) return (size_t)k;
  return ((size_t)il * p->n_xl + xl) * (size_t)p->n_tau + k;
}

/* Mask predicate of one contribution (aperture + angle), evaluated with the
 * mask velocity vm.  Monotone in tau when vm*tau is non-decreasing. */
static inline int mask_pass(const ctx_t* c, double hs, double hr, double zm) {
  const pstm_params_t* p = c->p;
  double ap = p->aperture_m;
  if (p->flags & PSTM_FLAG_TIME_APERTURE) {
    double a2 = zm * c->tan_th;
    if (a2 < ap) ap = a2;
  }
  if (!(hs < ap && hr < ap)) return 0;
  if (!(atan2(hs, zm + EPS_R) < c->th_rad)) return 0;
  if (!(atan2(hr, zm + EPS_R) < c->th_rad)) return 0;
  return 1;
}

/* Cosine edge taper of x towards its limit xmax
 * (processors/migration/aperture.py:226-270, same shape as weights.py:192-236):
 * 1 for x <= s = xmax (1 - tw), 0.5 (1 + cos(pi (x - s) / (xmax - s))) on
 * (s, xmax], 0 beyond xmax. */
static inline double edge_taper(double x, double xmax, double tw) {
  double s = xmax * (1.0 - tw);
  if (x > xmax) return 0.0;
  if (x > s) return 0.5 * (1.0 + cos(M_PI * (x - s) / (xmax - s)));
  return 1.0;
}

/* Offset window with near and far cosine tapers
 * (processors/migration/aperture.py:184-224); 1 when no window is set. */
static inline double offset_taper(const pstm_params_t* p, double off) {
  if (!(p->max_offset_m > 0)) return 1.0;
  double lo = p->min_offset_m, hi = p->max_offset_m;
  if (!(off >= lo && off <= hi)) return 0.0;
  double tr = p->taper_width * (hi - lo);
  double near = (off < lo + tr) ? 0.5 * (1.0 - cos(M_PI * (off - lo) / tr)) : 1.0;
  double far = (off > hi - tr) ? 0.5 * (1.0 + cos(M_PI * (off - (hi - tr)) / tr)) : 1.0;
  return near * far;
}

/* Inside the offset window of the plan (all offsets without one). */
static inline int offset_kept(const pstm_params_t* p, double off) {
  return !(p->max_offset_m > 0) || (off >= p->min_offset_m && off <= p->max_offset_m);
}

/* Taper composition of processors/migration/aperture.py:147-182: distance
 * tapers of both legs towards the aperture (the time-dependent aperture under
 * PSTM_FLAG_TIME_APERTURE), angle tapers of both legs with angle
 * atan2(h, |z|) towards max_angle, and the trace's offset taper. */
static inline double taper_value(const ctx_t* c, double hs, double hr, double zm, double off) {
  const pstm_params_t* p = c->p;
  double ap = p->aperture_m;
  if (p->flags & PSTM_FLAG_TIME_APERTURE) {
    double a2 = zm * c->tan_th;
    if (a2 < ap) ap = a2;
  }
  double tw = p->taper_width;
  double t = edge_taper(hs, ap, tw) * edge_taper(hr, ap, tw);
  t *= edge_taper(atan2(hs, fabs(zm)), c->th_rad, tw) * edge_taper(atan2(hr, fabs(zm)), c->th_rad, tw);
  return t * offset_taper(p, off);
}

/* traveltime_vti.py:164-222 with v0 = v_nmo = v (delta = 0), eta given. */
static inline double leg_time_vti(double h2, double z, double v, double eta) {
  double z_safe = fabs(z) > 1e-6 ? fabs(z) : 1e-6;
  double vn2 = v * v;
  double t0 = z_safe / v;
  double thyp = t0 * t0 + h2 / vn2;
  double den = vn2 * (t0 * t0 * vn2 + (1.0 + 2.0 * eta) * h2);
  if (den < 1e-10) den = 1e-10;
  double corr = 2.0 * eta * h2 * h2 / den;
  double t2 = thyp - corr;
  if (t2 < 1e-12) t2 = 1e-12;
  return sqrt(t2);
}

typedef struct {
  double w;        /* weight incl. taper */
  double u;        /* clamped sample position */
  double L;        /* antialias half-width in samples (1 without AA) */
  int clamped;     /* u was clamped */
  /* velocity derivatives (isotropic, no AA) */
  double dw_dv, du_dv;
} contrib_t;

/* Evaluate one (trace, image sample) pair.  Returns 0 when masked out. */
static int eval_pair(const ctx_t* c, double dxs, double dys, double dxr, double dyr,
                     double tau, double v, double vm, double eta, int want_grad,
                     contrib_t* o) {
  const pstm_params_t* p = c->p;
  double hs2 = dxs * dxs + dys * dys, hr2 = dxr * dxr + dyr * dyr;
  double hs = sqrt(hs2), hr = sqrt(hr2);
  double zm = vm * tau / 2.0;
  if (!mask_pass(c, hs, hr, zm)) return 0;
  double z = v * tau / 2.0;
  double rs = sqrt(hs2 + z * z), rr = sqrt(hr2 + z * z);
  double t;
  if (p->flags & PSTM_FLAG_VTI)
    t = leg_time_vti(hs2, z, v, eta) + leg_time_vti(hr2, z, v, eta);
  else
    t = (rs + rr) / v;
  double w;
  if (p->weight_mode == PSTM_WEIGHT_FULL)
    w = (z / (rs + EPS_R)) * (z / (rr + EPS_R)) / (rs * rr + EPS_R);
  else if (p->weight_mode == PSTM_WEIGHT_SPREADING)
    w = 1.0 / (rs * rr + EPS_R);
  else if (p->weight_mode == PSTM_WEIGHT_OBLIQUITY)
    w = (z / (rs + EPS_R)) * (z / (rr + EPS_R));
  else
    w = 1.0;
  double tap = 1.0;
  if (p->flags & PSTM_FLAG_TAPER) {
    tap = taper_value(c, hs, hr, zm, hypot(dxs - dxr, dys - dyr));
    w *= tap;
  }
  double n = p->n_samples;
  double u_raw = (t - p->t0) / p->dt;
  if ((p->flags & PSTM_FLAG_ZERO_OUTSIDE) && (u_raw < 0.0 || u_raw > n - 1.0)) return 0;
  double u = u_raw;
  o->clamped = 0;
  if (u < 0.0) { u = 0.0; o->clamped = 1; }
  if (u > n - 1.001) { u = n - 1.001; o->clamped = 1; }
  o->w = w;
  o->u = u;
  o->L = 1.0;
  if (p->flags & PSTM_FLAG_AA) {
    double ex = p->il_dx, ey = p->il_dy, fx = p->xl_dx, fy = p->xl_dy;
    /* d t / d(image position) per leg; a leg of zero length (tau = 0 and the
     * source or receiver on the column) has no dip: its term is 0. */
    double is = rs > 0.0 ? 1.0 / (v * rs) : 0.0, ir = rr > 0.0 ? 1.0 / (v * rr) : 0.0;
    double g_il = (dxs * ex + dys * ey) * is + (dxr * ex + dyr * ey) * ir;
    double g_xl = (dxs * fx + dys * fy) * is + (dxr * fx + dyr * fy) * ir;
    double L = p->aa_scale * (fabs(g_il) + fabs(g_xl)) / p->dt;
    if (L < 1.0) L = 1.0;
    if (L > p->aa_max) L = p->aa_max;
    o->L = L;
  }
  if (want_grad) {
    double dz = tau / 2.0;
    double drs = z * dz / rs, drr = z * dz / rr;
    double dt_dv = (drs + drr) / v - (rs + rr) / (v * v);
    double dw;
    if (p->weight_mode == PSTM_WEIGHT_FULL) {
      double D = (rs + EPS_R) * (rr + EPS_R) * (rs * rr + EPS_R);
      double dlogD = drs / (rs + EPS_R) + drr / (rr + EPS_R) + (drs * rr + rs * drr) / (rs * rr + EPS_R);
      double w0 = z * z / D;
      dw = 2.0 * z * dz / D - w0 * dlogD;
    } else if (p->weight_mode == PSTM_WEIGHT_SPREADING) {
      double den = rs * rr + EPS_R;
      dw = -(drs * rr + rs * drr) / (den * den);
    } else if (p->weight_mode == PSTM_WEIGHT_OBLIQUITY) {
      double D = (rs + EPS_R) * (rr + EPS_R);
      dw = 2.0 * z * dz / D - (z * z / D) * (drs / (rs + EPS_R) + drr / (rr + EPS_R));
    } else {
      dw = 0.0;
    }
    o->dw_dv = dw * tap;
    o->du_dv = o->clamped ? 0.0 : dt_dv / p->dt;
  }
  return 1;
}

static inline double lerp_at(const double* a, double x) {
  long i = (long)floor(x);
  double f = x - (double)i;
  return a[i] + f * (a[i + 1] - a[i]);
}

/* Double integral (causal, then anticausal) of the zero-extended trace.
 * out has n + 2*pad entries; out[e] belongs to sample e - pad. */
static void integrate_trace(const double* d, int n, int pad, double* out) {
  int ne = n + 2 * pad;
  double run = 0.0;
  for (int e = 0; e < ne; ++e) {
    int i = e - pad;
    if (i >= 0 && i < n) run += d[i];
    out[e] = run;
  }
  double acc = 0.0;
  for (int e = ne - 1; e >= 0; --e) {
    acc += out[e];
    out[e] = acc;
  }
}

/* Amplitude of one contribution: linear interpolation (A4) or the triangle. */
static inline double sample_amp(const ctx_t* c, const double* d, const double* aint,
                                const contrib_t* o, double* slope) {
  const pstm_params_t* p = c->p;
  if (p->flags & PSTM_FLAG_AA) {
    double x = o->u + c->pad;
    double L = o->L;
    if (slope) *slope = 0.0;
    return (2.0 * lerp_at(aint, x) - lerp_at(aint, x - L) - lerp_at(aint, x + L)) / (L * L);
  }
  long i0 = (long)o->u;   /* u >= 0: truncation is floor, kirchhoff_kernel.py:716 */
  long i1 = i0 + 1;
  if (i1 > p->n_samples - 1) i1 = p->n_samples - 1;
  double f = o->u - (double)i0;
  if (slope) *slope = d[i1] - d[i0];
  return d[i0] + f * (d[i1] - d[i0]);
}

/* First tau index whose mask passes, or n_tau. */
static int first_pass(const ctx_t* c, const double* vm_col, int vm_stride_is_1d,
                      double hs, double hr, int mono) {
  const pstm_params_t* p = c->p;
  int nt = p->n_tau;
  (void)vm_stride_is_1d;
  if (!mono || g_exhaustive) return -1;  /* caller evaluates every k */
  int lo = 0, hi = nt;                     /* predicate false on [0, lo), true on [hi, nt) */
  while (lo < hi) {
    int mid = lo + (hi - lo) / 2;
    double tau = p->tau0 + mid * p->dtau;
    if (mask_pass(c, hs, hr, vm_col[mid] * tau / 2.0)) hi = mid; else lo = mid + 1;
  }
  return lo;
}

static int column_monotone(const pstm_params_t* p, const double* vm_col) {
  double prev = -INFINITY;
  for (int k = 0; k < p->n_tau; ++k) {
    double z = vm_col[k] * (p->tau0 + k * p->dtau);
    if (z < prev) return 0;
    prev = z;
  }
  return p->tau0 >= 0.0;
}

static void column_xy(const pstm_params_t* p, int il, int xl, double* x, double* y) {
  *x = p->origin_x + il * p->il_dx + xl * p->xl_dx;
  *y = p->origin_y + il * p->il_dy + xl * p->xl_dy;
}

/* Gather velocity (or eta) of one column into a contiguous buffer. */
static void column_vals(const pstm_params_t* p, const double* src, int il, int xl, double* out) {
  for (int k = 0; k < p->n_tau; ++k) out[k] = src[vidx(p, il, xl, k)];
}

/*
 * Forward: image[il][xl][k] += sum over traces of w * amp; fold += count.
 * traces: [n_traces][n_samples] FP64.  sx..ry: world coordinates.
 * vel/eta: layout p->vel_mode.  mask_vel: velocity used by the masks and
 * taper (NULL = vel).  image/fold: [il][xl][tau] FP64 (fold may be NULL).
 */
int oracle_forward(const pstm_params_t* p, const double* traces, const double* sx,
                   const double* sy, const double* rx, const double* ry, const double* vel,
                   const double* eta, const double* mask_vel, double* image, double* fold) {
  ctx_t c;
  ctx_init(&c, p);
  const int n = p->n_samples, nt = p->n_tau;
  const int64_t ntr = p->n_traces;
  const int do_aa = (p->flags & PSTM_FLAG_AA) != 0;
  const int ne = n + 2 * c.pad;
  double* aint = NULL;
  if (do_aa) {
    aint = (double*)malloc(sizeof(double) * (size_t)ntr * ne);
    if (!aint) return -1;
#pragma omp parallel for schedule(static)
    for (int64_t j = 0; j < ntr; ++j) integrate_trace(traces + j * n, n, c.pad, aint + j * ne);
  }
  const int ncol = p->n_il * p->n_xl;
#pragma omp parallel
  {
    double* vcol = (double*)malloc(sizeof(double) * nt);
    double* mcol = (double*)malloc(sizeof(double) * nt);
    double* ecol = (double*)malloc(sizeof(double) * nt);
#pragma omp for schedule(dynamic, 1)
    for (int col = 0; col < ncol; ++col) {
      int il = col / p->n_xl, xl = col % p->n_xl;
      double x, y;
      column_xy(p, il, xl, &x, &y);
      column_vals(p, vel, il, xl, vcol);
      column_vals(p, mask_vel ? mask_vel : vel, il, xl, mcol);
      if ((p->flags & PSTM_FLAG_VTI) && eta) column_vals(p, eta, il, xl, ecol);
      else memset(ecol, 0, sizeof(double) * nt);
      int mono = column_monotone(p, mcol);
      double* img = image + (size_t)col * nt;
      double* fld = fold ? fold + (size_t)col * nt : NULL;
      for (int64_t j = 0; j < ntr; ++j) {
        if (!offset_kept(p, hypot(rx[j] - sx[j], ry[j] - sy[j]))) continue;
        double dxs = x - sx[j], dys = y - sy[j], dxr = x - rx[j], dyr = y - ry[j];
        double hs = sqrt(dxs * dxs + dys * dys), hr = sqrt(dxr * dxr + dyr * dyr);
        if (!(hs < p->aperture_m && hr < p->aperture_m)) continue;
        int k0 = first_pass(&c, mcol, 0, hs, hr, mono);
        const double* d = traces + j * n;
        const double* a = do_aa ? aint + j * ne : NULL;
        for (int k = (k0 < 0 ? 0 : k0); k < nt; ++k) {
          double tau = p->tau0 + k * p->dtau;
          contrib_t o;
          if (!eval_pair(&c, dxs, dys, dxr, dyr, tau, vcol[k], mcol[k], ecol[k], 0, &o)) continue;
          img[k] += o.w * sample_amp(&c, d, a, &o, NULL);
          if (fld) fld[k] += 1.0;
        }
      }
    }
    free(vcol);
    free(mcol);
    free(ecol);
  }
  free(aint);
  return 0;
}

/*
 * Adjoint: traces_out[j][i] += (L^T image)[j][i], the exact transpose of
 * oracle_forward (same masks, weights, interpolation, antialias taps).
 */
int oracle_adjoint(const pstm_params_t* p, const double* image, const double* sx,
                   const double* sy, const double* rx, const double* ry, const double* vel,
                   const double* eta, double* traces_out) {
  ctx_t c;
  ctx_init(&c, p);
  const int n = p->n_samples, nt = p->n_tau;
  const int64_t ntr = p->n_traces;
  const int do_aa = (p->flags & PSTM_FLAG_AA) != 0;
  const int ne = n + 2 * c.pad;
  const int ncol = p->n_il * p->n_xl;
#pragma omp parallel
  {
    double* vcol = (double*)malloc(sizeof(double) * nt);
    double* ecol = (double*)malloc(sizeof(double) * nt);
    double* abar = (double*)malloc(sizeof(double) * (do_aa ? ne : 1));
#pragma omp for schedule(dynamic, 1)
    for (int64_t j = 0; j < ntr; ++j) {
      double* dout = traces_out + j * n;
      if (!offset_kept(p, hypot(rx[j] - sx[j], ry[j] - sy[j]))) continue;
      if (do_aa) memset(abar, 0, sizeof(double) * ne);
      for (int col = 0; col < ncol; ++col) {
        int il = col / p->n_xl, xl = col % p->n_xl;
        double x, y;
        column_xy(p, il, xl, &x, &y);
        double dxs = x - sx[j], dys = y - sy[j], dxr = x - rx[j], dyr = y - ry[j];
        double hs = sqrt(dxs * dxs + dys * dys), hr = sqrt(dxr * dxr + dyr * dyr);
        if (!(hs < p->aperture_m && hr < p->aperture_m)) continue;
        column_vals(p, vel, il, xl, vcol);
        if ((p->flags & PSTM_FLAG_VTI) && eta) column_vals(p, eta, il, xl, ecol);
        else memset(ecol, 0, sizeof(double) * nt);
        int mono = column_monotone(p, vcol);
        int k0 = first_pass(&c, vcol, 0, hs, hr, mono);
        const double* m = image + (size_t)col * nt;
        for (int k = (k0 < 0 ? 0 : k0); k < nt; ++k) {
          double tau = p->tau0 + k * p->dtau;
          contrib_t o;
          if (!eval_pair(&c, dxs, dys, dxr, dyr, tau, vcol[k], vcol[k], ecol[k], 0, &o)) continue;
          double g = o.w * m[k];
          if (do_aa) {
            double L = o.L, x0 = o.u + c.pad;
            double cc = 2.0 / (L * L), ce = -1.0 / (L * L);
            double xs[3] = {x0, x0 - L, x0 + L};
            double cs[3] = {cc, ce, ce};
            for (int q = 0; q < 3; ++q) {
              long i = (long)floor(xs[q]);
              double f = xs[q] - (double)i;
              abar[i] += g * cs[q] * (1.0 - f);
              abar[i + 1] += g * cs[q] * f;
            }
          } else {
            long i0 = (long)o.u;
            long i1 = i0 + 1 > n - 1 ? n - 1 : i0 + 1;
            double f = o.u - (double)i0;
            dout[i0] += g * (1.0 - f);
            dout[i1] += g * f;
          }
        }
      }
      if (do_aa) {
        /* The double-integration operator (causal then anticausal) is
         * symmetric, so its transpose is itself. */
        double run = 0.0;
        for (int e = 0; e < ne; ++e) { run += abar[e]; abar[e] = run; }
        double acc = 0.0;
        for (int e = ne - 1; e >= 0; --e) { acc += abar[e]; abar[e] = acc; }
        for (int i = 0; i < n; ++i) dout[i] += abar[i + c.pad];
      }
    }
    free(vcol);
    free(ecol);
    free(abar);
  }
  return 0;
}

/*
 * Velocity gradient: vgrad[il][xl][k] += image_grad[il][xl][k] *
 * d(image[il][xl][k]) / d(v[il][xl][k]) with masks and taper held at the
 * same velocity (treated as constants).  Isotropic, no antialias, no
 * zero-outside; vel must be 3D.
 */
int oracle_vel_grad(const pstm_params_t* p, const double* traces, const double* sx,
                    const double* sy, const double* rx, const double* ry, const double* vel,
                    const double* image_grad, double* vgrad) {
  if (p->flags & (PSTM_FLAG_AA | PSTM_FLAG_VTI | PSTM_FLAG_ZERO_OUTSIDE)) return -2;
  if (p->vel_mode != PSTM_VEL_3D) return -2;
  ctx_t c;
  ctx_init(&c, p);
  const int n = p->n_samples, nt = p->n_tau;
  const int64_t ntr = p->n_traces;
  const int ncol = p->n_il * p->n_xl;
#pragma omp parallel
  {
    double* vcol = (double*)malloc(sizeof(double) * nt);
#pragma omp for schedule(dynamic, 1)
    for (int col = 0; col < ncol; ++col) {
      int il = col / p->n_xl, xl = col % p->n_xl;
      double x, y;
      column_xy(p, il, xl, &x, &y);
      column_vals(p, vel, il, xl, vcol);
      int mono = column_monotone(p, vcol);
      const double* g = image_grad + (size_t)col * nt;
      double* out = vgrad + (size_t)col * nt;
      for (int64_t j = 0; j < ntr; ++j) {
        if (!offset_kept(p, hypot(rx[j] - sx[j], ry[j] - sy[j]))) continue;
        double dxs = x - sx[j], dys = y - sy[j], dxr = x - rx[j], dyr = y - ry[j];
        double hs = sqrt(dxs * dxs + dys * dys), hr = sqrt(dxr * dxr + dyr * dyr);
        if (!(hs < p->aperture_m && hr < p->aperture_m)) continue;
        int k0 = first_pass(&c, vcol, 0, hs, hr, mono);
        const double* d = traces + j * n;
        for (int k = (k0 < 0 ? 0 : k0); k < nt; ++k) {
          double tau = p->tau0 + k * p->dtau;
          contrib_t o;
          if (!eval_pair(&c, dxs, dys, dxr, dyr, tau, vcol[k], vcol[k], 0.0, 1, &o)) continue;
          double slope;
          double a = sample_amp(&c, d, NULL, &o, &slope);
          out[k] += g[k] * (o.dw_dv * a + o.w * slope * o.du_dv);
        }
      }
    }
    free(vcol);
  }
  return 0;
}

/* Single-pair helpers exposed for known-answer tests. */
double oracle_leg_time_vti(double h2, double z, double v, double eta) {
  return leg_time_vti(h2, z, v, eta);
}

int oracle_pair(const pstm_params_t* p, double dxs, double dys, double dxr, double dyr,
                double tau, double v, double eta, double* w, double* u, double* L) {
  ctx_t c;
  ctx_init(&c, p);
  contrib_t o;
  int ok = eval_pair(&c, dxs, dys, dxr, dyr, tau, v, v, eta, 0, &o);
  if (ok) { *w = o.w; *u = o.u; *L = o.L; }
  return ok;
}
