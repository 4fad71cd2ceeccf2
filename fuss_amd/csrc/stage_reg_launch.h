// stage_reg_launch.h -- launcher of the register-column stage kernel
// (template definition; stage_reg_{d,f}.hip instantiate it).
#pragma once

#include "stage_launch.h"
#include "fuss_stage_reg.h"

namespace fuss {

template <bool UM, bool QU, int KCT, typename R>
void launch_stage3r_k(fuss_ctx* c, const StageArgs<R>& a, unsigned ntile, bool prim, bool irs) {
  const unsigned g = (ntile + RC_WAVES - 1) / RC_WAVES;
  constexpr unsigned T = 64 * RC_WAVES;
  const int nt = (int)ntile;
  if (prim && a.dtw) {  // first stage of a steady step: dt here
    irs ? k_stage3r<UM, QU, true, true, true, false, R, KCT><<<g, T, 0, c->stream>>>(a, nt)
        : k_stage3r<UM, QU, true, false, true, false, R, KCT><<<g, T, 0, c->stream>>>(a, nt);
  } else if (a.use_dtg) {  // time-accurate: the global dt
    if (prim) irs ? k_stage3r<UM, QU, true, true, false, true, R, KCT><<<g, T, 0, c->stream>>>(a, nt)
                  : k_stage3r<UM, QU, true, false, false, true, R, KCT><<<g, T, 0, c->stream>>>(a, nt);
    else irs ? k_stage3r<UM, QU, false, true, false, true, R, KCT><<<g, T, 0, c->stream>>>(a, nt)
             : k_stage3r<UM, QU, false, false, false, true, R, KCT><<<g, T, 0, c->stream>>>(a, nt);
  } else {
    if (prim) irs ? k_stage3r<UM, QU, true, true, false, false, R, KCT><<<g, T, 0, c->stream>>>(a, nt)
                  : k_stage3r<UM, QU, true, false, false, false, R, KCT><<<g, T, 0, c->stream>>>(a, nt);
    else irs ? k_stage3r<UM, QU, false, true, false, false, R, KCT><<<g, T, 0, c->stream>>>(a, nt)
             : k_stage3r<UM, QU, false, false, false, false, R, KCT><<<g, T, 0, c->stream>>>(a, nt);
  }
}

template <int KCT, typename R>
void launch_stage3r_u(fuss_ctx* c, Level& L, const StageArgs<R>& a, bool prim, bool irs) {
  StageArgs<R> at = a;
  at.ctiles = L.ctiles.p;
  const unsigned n = (unsigned)L.hctiles.size();
  if (L.all_umat)
    L.all_uqv ? launch_stage3r_k<true, true, KCT>(c, at, n, prim, irs) : launch_stage3r_k<true, false, KCT>(c, at, n, prim, irs);
  else
    L.all_uqv ? launch_stage3r_k<false, true, KCT>(c, at, n, prim, irs)
              : launch_stage3r_k<false, false, KCT>(c, at, n, prim, irs);
}

// The register-column kernel on the level's 4-row columns (rectilinear
// levels: a.rect set; tables starting at T >= 0; column heights 8, 16, 32
// as rc_column_ok)
template <typename R>
void launch_stage3r(fuss_ctx* c, Level& L, const StageArgs<R>& a, bool prim, bool irs) {
  if (L.kc == 8) launch_stage3r_u<8>(c, L, a, prim, irs);
  else if (L.kc == 16) launch_stage3r_u<16>(c, L, a, prim, irs);
  else launch_stage3r_u<32>(c, L, a, prim, irs);
}

}  // namespace fuss
