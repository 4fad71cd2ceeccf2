// stage_launch.h -- launchers of the stage kernels (template definitions;
// each stage_*.hip unit instantiates its share for double and float).
#pragma once

#include "fuss_host.h"
#include "fuss_stage.h"

namespace fuss {

template <int NDIR, bool TANG, bool PRIM, bool IRSMODE, typename R>
void launch_stage_t(fuss_ctx* c, Level& L, const StageArgs<R>& a) {
  k_stage<NDIR, TANG, PRIM, IRSMODE, R><<<(unsigned)L.htiles.size(), TILE_T, 0, c->stream>>>(a);
}

template <int NDIR, typename R>
void launch_stage_n(fuss_ctx* c, Level& L, const StageArgs<R>& a, bool tang, bool prim, bool irs) {
  if (tang) {
    if (prim) { irs ? launch_stage_t<NDIR, true, true, true>(c, L, a) : launch_stage_t<NDIR, true, true, false>(c, L, a); }
    else { irs ? launch_stage_t<NDIR, true, false, true>(c, L, a) : launch_stage_t<NDIR, true, false, false>(c, L, a); }
  } else {
    if (prim) { irs ? launch_stage_t<NDIR, false, true, true>(c, L, a) : launch_stage_t<NDIR, false, true, false>(c, L, a); }
    else { irs ? launch_stage_t<NDIR, false, false, true>(c, L, a) : launch_stage_t<NDIR, false, false, false>(c, L, a); }
  }
}

template <typename R>
void launch_stage2d(fuss_ctx* c, Level& L, const StageArgs<R>& a, bool tang, bool prim, bool irs) {
  launch_stage_n<2>(c, L, a, tang, prim, irs);
}

template <bool UM, bool RECT, int KCT, int TY, typename R>
void launch_stage3p(fuss_ctx* c, unsigned g, const StageArgs<R>& a, bool prim, bool irs) {
  constexpr int T = TILE_X * TY;
  if (RECT && prim && a.dtw) {  // first stage of a steady step: dt here (rectilinear, primitive)
    irs ? k_stage3p<UM, true, true, true, true, R, KCT, TY><<<g, T, 0, c->stream>>>(a)
        : k_stage3p<UM, true, false, true, true, R, KCT, TY><<<g, T, 0, c->stream>>>(a);
  } else if (prim) {
    irs ? k_stage3p<UM, true, true, RECT, false, R, KCT, TY><<<g, T, 0, c->stream>>>(a)
        : k_stage3p<UM, true, false, RECT, false, R, KCT, TY><<<g, T, 0, c->stream>>>(a);
  } else {
    irs ? k_stage3p<UM, false, true, RECT, false, R, KCT, TY><<<g, T, 0, c->stream>>>(a)
        : k_stage3p<UM, false, false, RECT, false, R, KCT, TY><<<g, T, 0, c->stream>>>(a);
  }
}

// the LDS-ring kernel on the level's 4-row (ctiles) or 8-row (c8tiles) columns
template <int KCT, int TY, typename R>
void launch_stage3p_t(fuss_ctx* c, Level& L, const StageArgs<R>& a, bool prim, bool irs) {
  StageArgs<R> at = a;
  at.ctiles = TY == 8 ? L.c8tiles.p : L.ctiles.p;
  const unsigned g = (unsigned)(TY == 8 ? L.hc8tiles.size() : L.hctiles.size());
  if (L.all_umat)
    a.rect ? launch_stage3p<true, true, KCT, TY>(c, g, at, prim, irs)
           : launch_stage3p<true, false, KCT, TY>(c, g, at, prim, irs);
  else
    a.rect ? launch_stage3p<false, true, KCT, TY>(c, g, at, prim, irs)
           : launch_stage3p<false, false, KCT, TY>(c, g, at, prim, irs);
}

template <typename R, int TY>
void launch_stage3p_ring(fuss_ctx* c, Level& L, const StageArgs<R>& a, bool prim, bool irs) {
  if (L.kc == 4) launch_stage3p_t<4, TY>(c, L, a, prim, irs);
  else if (L.kc == 16) launch_stage3p_t<16, TY>(c, L, a, prim, irs);
  else if (L.kc == 32) launch_stage3p_t<32, TY>(c, L, a, prim, irs);
  else if (L.kc == 64) launch_stage3p_t<64, TY>(c, L, a, prim, irs);
  else launch_stage3p_t<8, TY>(c, L, a, prim, irs);
}

// blocks with tangential weights: the pipelined kernel on the 4-row columns
template <bool UM, int KCT, typename R>
void launch_stage3t(fuss_ctx* c, unsigned g, const StageArgs<R>& a, bool prim, bool irs) {
  if (prim) { irs ? k_stage3t<UM, true, true, R, KCT><<<g, TILE_T, 0, c->stream>>>(a) : k_stage3t<UM, true, false, R, KCT><<<g, TILE_T, 0, c->stream>>>(a); }
  else { irs ? k_stage3t<UM, false, true, R, KCT><<<g, TILE_T, 0, c->stream>>>(a) : k_stage3t<UM, false, false, R, KCT><<<g, TILE_T, 0, c->stream>>>(a); }
}

template <int KCT, typename R>
void launch_stage3d_k(fuss_ctx* c, Level& L, const StageArgs<R>& a, bool tang, bool prim, bool irs) {
  const unsigned g = (unsigned)L.hctiles.size();
  const bool row = L.row_march >= 0 ? L.row_march == 1 : !c->stage_lds;
  if (tang) {
    L.all_umat ? launch_stage3t<true, KCT>(c, g, a, prim, irs) : launch_stage3t<false, KCT>(c, g, a, prim, irs);
  } else if (!row && !L.dims2) {
    if (prim) { irs ? k_stage3d<false, true, true, R, KCT><<<g, TILE_T, 0, c->stream>>>(a) : k_stage3d<false, true, false, R, KCT><<<g, TILE_T, 0, c->stream>>>(a); }
    else { irs ? k_stage3d<false, false, true, R, KCT><<<g, TILE_T, 0, c->stream>>>(a) : k_stage3d<false, false, false, R, KCT><<<g, TILE_T, 0, c->stream>>>(a); }
  } else if (!row) {
    // (the LDS-ring kernel dispatches on L.kc itself)
  } else {
    if (prim) { irs ? k_stage3w<true, true, R, KCT><<<g, TILE_T, 0, c->stream>>>(a) : k_stage3w<true, false, R, KCT><<<g, TILE_T, 0, c->stream>>>(a); }
    else { irs ? k_stage3w<false, true, R, KCT><<<g, TILE_T, 0, c->stream>>>(a) : k_stage3w<false, false, R, KCT><<<g, TILE_T, 0, c->stream>>>(a); }
  }
}

// the register-column kernel (stage_reg_{d,f}.hip)
template <typename R>
void launch_stage3r(fuss_ctx* c, Level& L, const StageArgs<R>& a, bool prim, bool irs);

// Levels the register-column kernel runs: per-axis metric factors in use,
// table rows from T >= 0 (its row conversion saturates at 0), a column
// height it is built for.
template <typename R>
inline bool rc_eligible(const Level& L, const StageArgs<R>& a) {
  return a.rect && a.tab.Ti1 >= 0 && (L.kc == 8 || L.kc == 16 || L.kc == 32);
}

template <typename R>
void launch_stage3d(fuss_ctx* c, Level& L, const StageArgs<R>& a, bool tang, bool prim, bool irs) {
  const bool row = L.row_march >= 0 ? L.row_march == 1 : !c->stage_lds;
  if (!tang && !row && L.dims2) {
    if (L.stage_ty == FUSS_STAGE_REGCOL && rc_eligible(L, a)) launch_stage3r<R>(c, L, a, prim, irs);
    else if (L.stage_ty == 8) launch_stage3p_ring<R, 8>(c, L, a, prim, irs);
    else launch_stage3p_ring<R, 4>(c, L, a, prim, irs);
    return;
  }
  if (L.kc == 4) launch_stage3d_k<4>(c, L, a, tang, prim, irs);
  else if (L.kc == 16) launch_stage3d_k<16>(c, L, a, tang, prim, irs);
  else if (L.kc == 32) launch_stage3d_k<32>(c, L, a, tang, prim, irs);
  else if (L.kc == 64) launch_stage3d_k<64>(c, L, a, tang, prim, irs);
  else launch_stage3d_k<8>(c, L, a, tang, prim, irs);
}

// External_Function path: Fluxes + BC_Fluxes to R, then RK_Newstate
template <typename R>
void launch_stage_raw(fuss_ctx* c, Level& L, const StageArgs<R>& a, int ndir, bool tang) {
  const unsigned gt = (unsigned)L.htiles.size();
  if (ndir == 3)
    tang ? k_stage_raw<3, true, R><<<gt, TILE_T, 0, c->stream>>>(a) : k_stage_raw<3, false, R><<<gt, TILE_T, 0, c->stream>>>(a);
  else
    tang ? k_stage_raw<2, true, R><<<gt, TILE_T, 0, c->stream>>>(a) : k_stage_raw<2, false, R><<<gt, TILE_T, 0, c->stream>>>(a);
}

template <typename R>
void launch_newstate(fuss_ctx* c, Level& L, const StageArgs<R>& a, bool prim, bool irs) {
  const unsigned gt = (unsigned)L.htiles.size();
  if (prim)
    irs ? k_newstate<true, true, R><<<gt, TILE_T, 0, c->stream>>>(a) : k_newstate<true, false, R><<<gt, TILE_T, 0, c->stream>>>(a);
  else
    irs ? k_newstate<false, true, R><<<gt, TILE_T, 0, c->stream>>>(a) : k_newstate<false, false, R><<<gt, TILE_T, 0, c->stream>>>(a);
}

}  // namespace fuss
