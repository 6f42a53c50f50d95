// pstm_adj.hip -- adjoint (demigration) tile kernels and the dispatch of all
// three tile operators.
#include "pstm_tile.h"

namespace pstm {

int launch_forward(const pstm_params_t* p, const PlanLayout& L, const TileArgs& a, hipStream_t s) {
  int rc;
  if (p->dtype == PSTM_F32) rc = launch_forward_dt<PSTM_F32>(p, L, a, s);
  else if (p->dtype == PSTM_BF16) rc = launch_forward_dt<PSTM_BF16>(p, L, a, s);
  else rc = launch_forward_dt<PSTM_F16>(p, L, a, s);
  return rc != PSTM_OK ? rc : check_launch("forward");
}

int launch_adjoint(const pstm_params_t* p, const PlanLayout& L, const TileArgs& a, hipStream_t s) {
  const bool aa = (p->flags & PSTM_FLAG_AA) != 0, vti = (p->flags & PSTM_FLAG_VTI) != 0;
  const bool gen = !fast_path(p);
  const int64_t win = whole_window(p, a.aa_pad);
  int rc;
  // Windows accumulate in 64-bit fixed point (integer LDS atomics) whenever
  // a whole trace window fits at 8 bytes per sample -- always with antialias,
  // whose FP64 windows have that size; longer plain windows fall back to
  // FP32 windows and float LDS atomics.
#define PSTM_ADJ(TX, TY, AAV, VTIV, GENV, OCCV, FXV) \
  launch_one<MODE_ADJ, TX, TY, PSTM_F32, AAV, VTIV, GENV, OCCV, FXV>(a, s)
#define PSTM_ADJ4(TX, TY, GENV, OCCV, FXV)                            \
  (aa ? (vti ? PSTM_ADJ(TX, TY, true, true, GENV, 2, true)            \
             : PSTM_ADJ(TX, TY, true, false, GENV, 2, true))          \
      : (vti ? PSTM_ADJ(TX, TY, false, true, GENV, OCCV, FXV)         \
             : PSTM_ADJ(TX, TY, false, false, GENV, OCCV, FXV)))
  if (long_traces(p, a)) {
    // windows may exceed every form's LDS share: general forms at the lowest
    // occupancy (FP32 windows plain, fixed point with antialias), staging
    // such windows in segments
#define PSTM_ADJS(TX, TY)                                                                                          \
  (aa ? (vti ? launch_one<MODE_ADJ, TX, TY, PSTM_F32, true, true, true, 2, true, false, PSTM_WEIGHT_FULL, true>(a, s)    \
             : launch_one<MODE_ADJ, TX, TY, PSTM_F32, true, false, true, 2, true, false, PSTM_WEIGHT_FULL, true>(a, s))   \
      : (vti ? launch_one<MODE_ADJ, TX, TY, PSTM_F32, false, true, true, 3, false, false, PSTM_WEIGHT_FULL, true>(a, s)  \
             : launch_one<MODE_ADJ, TX, TY, PSTM_F32, false, false, true, 3, false, false, PSTM_WEIGHT_FULL, true>(a, s)))
    if (a.tile_x == 4 && a.tile_y == 2) rc = PSTM_ADJS(4, 2);
    else if (a.tile_x == 8 && a.tile_y == 1) rc = PSTM_ADJS(8, 1);
    else if (a.tile_x == 1 && a.tile_y == 8) rc = PSTM_ADJS(1, 8);
    else rc = set_error(PSTM_ERR_INVALID, "unsupported tile shape %dx%d", a.tile_x, a.tile_y);
#undef PSTM_ADJS
    return rc != PSTM_OK ? rc : check_launch("adjoint");
  }
  // plain fixed-point windows are bank-skewed (skew_len elements)
  const int64_t fxwin = aa ? win : skew_len((int)win);
  const bool fx8 = fxwin <= window_cap_elems<8, false, 3, true>();
  // Plain 3D forms: occupancy first (measured on the 4-class LSM slice: 4 x 2
  // tiles at 7 workgroups per CU 831 Gcontrib/s, at 4 752, 8 x 4 tiles at 2
  // 717; with a laterally varying velocity 4 x 2 tiles at 6 per CU 803, at 5
  // 780, at 4 735); longer traces fall back to the wider tiles' larger windows.
  if (a.tile_x == 4 && a.tile_y == 2 && !aa && !gen && !vti && p->vel_mode == PSTM_VEL_1D &&
      fxwin <= window_cap_elems<8, false, 7, true>()) {
    rc = launch_one<MODE_ADJ, 4, 2, PSTM_F32, false, false, false, 7, true, true>(a, s);
  } else if (a.tile_x == 4 && a.tile_y == 2 && !aa && !gen && !vti && fxwin <= window_cap_elems<8, false, 6, true>()) {
    rc = launch_one<MODE_ADJ, 4, 2, PSTM_F32, false, false, false, 6, true>(a, s);
  } else if (a.tile_x == 4 && a.tile_y == 2 && !aa && !gen && !vti && fxwin <= window_cap_elems<8, false, 4, true>()) {
    rc = launch_one<MODE_ADJ, 4, 2, PSTM_F32, false, false, false, 4, true>(a, s);
  } else if (a.tile_x == 4 && a.tile_y == 2 && !aa && !gen && !vti &&
      fxwin <= window_cap_elems<32, false, 2, true>()) {
    // 3D, isotropic, FULL weight, no antialias: 8 x 4-column tiles, so that
    // each staged trace window (flushed to the traces with global atomics)
    // serves 32 columns (the general and VTI forms keep 4 x 2 tiles: with 32
    // columns their registers spill)
    TileArgs b = a;
    b.tile_x = 8;
    b.tile_y = 4;
    b.n_tiles_x = (p->n_il + 7) / 8;
    b.n_tiles_y = (p->n_xl + 3) / 4;
    rc = p->vel_mode == PSTM_VEL_1D ? launch_one<MODE_ADJ, 8, 4, PSTM_F32, false, false, false, 2, true, true>(b, s)
                                    : launch_one<MODE_ADJ, 8, 4, PSTM_F32, false, false, false, 2, true>(b, s);
  } else if (a.tile_x == 4 && a.tile_y == 2) {
    rc = fx8 ? (gen ? PSTM_ADJ4(4, 2, true, 3, true) : PSTM_ADJ4(4, 2, false, 3, true))
             : (gen ? PSTM_ADJ4(4, 2, true, 3, false) : PSTM_ADJ4(4, 2, false, 3, false));
  } else if (a.tile_x == 8 && a.tile_y == 1) {
    rc = fx8 ? PSTM_ADJ4(8, 1, true, 3, true) : PSTM_ADJ4(8, 1, true, 3, false);
  } else if (a.tile_x == 1 && a.tile_y == 8) {
    rc = fx8 ? PSTM_ADJ4(1, 8, true, 3, true) : PSTM_ADJ4(1, 8, true, 3, false);
  } else {
    rc = set_error(PSTM_ERR_INVALID, "unsupported tile shape %dx%d", a.tile_x, a.tile_y);
  }
#undef PSTM_ADJ4
#undef PSTM_ADJ
  return rc != PSTM_OK ? rc : check_launch("adjoint");
}

}  // namespace pstm
