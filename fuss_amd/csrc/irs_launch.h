// irs_launch.h -- launchers of the IRS kernels (template definitions; each
// irs_*.hip unit instantiates its share).
#pragma once

#include <utility>

#include "fuss_host.h"
#include "fuss_irs.h"

namespace fuss {

template <typename R>
void launch_irs_z(fuss_ctx* c, Level& L, const IrsArgs<R>& ia, bool prim) {
  const unsigned g = (unsigned)L.hctiles.size();
  if (L.kc == 4) prim ? k_irs_z<true, R, 4><<<g, TILE_T, 0, c->stream>>>(ia) : k_irs_z<false, R, 4><<<g, TILE_T, 0, c->stream>>>(ia);
  else if (L.kc == 16) prim ? k_irs_z<true, R, 16><<<g, TILE_T, 0, c->stream>>>(ia) : k_irs_z<false, R, 16><<<g, TILE_T, 0, c->stream>>>(ia);
  else if (L.kc == 32) prim ? k_irs_z<true, R, 32><<<g, TILE_T, 0, c->stream>>>(ia) : k_irs_z<false, R, 32><<<g, TILE_T, 0, c->stream>>>(ia);
  else if (L.kc == 64) prim ? k_irs_z<true, R, 64><<<g, TILE_T, 0, c->stream>>>(ia) : k_irs_z<false, R, 64><<<g, TILE_T, 0, c->stream>>>(ia);
  else prim ? k_irs_z<true, R, 8><<<g, TILE_T, 0, c->stream>>>(ia) : k_irs_z<false, R, 8><<<g, TILE_T, 0, c->stream>>>(ia);
}

template <bool XCON, bool P2, bool DT, int TY, typename R>
void launch_irs_xyz_x(fuss_ctx* c, Level& L, const IrsArgs<R>& ia, bool prim) {
  const unsigned g = (unsigned)(TY == 8 ? L.hc8tiles.size() : L.hctiles.size());
  constexpr int T = Xyz<TY>::TT;
  if (L.kc == 4) prim ? k_irs_xyz<true, XCON, P2, DT, R, 4, TY><<<g, T, 0, c->stream>>>(ia) : k_irs_xyz<false, XCON, P2, DT, R, 4, TY><<<g, T, 0, c->stream>>>(ia);
  else if (L.kc == 16) prim ? k_irs_xyz<true, XCON, P2, DT, R, 16, TY><<<g, T, 0, c->stream>>>(ia) : k_irs_xyz<false, XCON, P2, DT, R, 16, TY><<<g, T, 0, c->stream>>>(ia);
  else if (L.kc == 32) prim ? k_irs_xyz<true, XCON, P2, DT, R, 32, TY><<<g, T, 0, c->stream>>>(ia) : k_irs_xyz<false, XCON, P2, DT, R, 32, TY><<<g, T, 0, c->stream>>>(ia);
  else if (L.kc == 64) prim ? k_irs_xyz<true, XCON, P2, DT, R, 64, TY><<<g, T, 0, c->stream>>>(ia) : k_irs_xyz<false, XCON, P2, DT, R, 64, TY><<<g, T, 0, c->stream>>>(ia);
  else prim ? k_irs_xyz<true, XCON, P2, DT, R, 8, TY><<<g, T, 0, c->stream>>>(ia) : k_irs_xyz<false, XCON, P2, DT, R, 8, TY><<<g, T, 0, c->stream>>>(ia);
}

template <bool XCON, bool P2, int TY, typename R>
void launch_irs_xyz_p(fuss_ctx* c, Level& L, const IrsArgs<R>& ia, bool prim) {
  ia.dtnext ? launch_irs_xyz_x<XCON, P2, true, TY>(c, L, ia, prim)
            : launch_irs_xyz_x<XCON, P2, false, TY>(c, L, ia, prim);
}

// P2: 1 + 2 beta is a power of two (beta = 0.5: 2), the sweeps multiply by
// its exact reciprocal; DT: the pass also computes the next step's dt.  The
// caller passes the tile list of L.xyz_ty (ctiles for 4 rows, c8tiles for 8).
template <typename R, int TY>
void launch_irs_xyz_ty(fuss_ctx* c, Level& L, const IrsArgs<R>& ia, bool prim) {
  const bool p2 = ia.inv != 0.0;
  if (L.no_xcon)
    p2 ? launch_irs_xyz_p<false, true, TY>(c, L, ia, prim) : launch_irs_xyz_p<false, false, TY>(c, L, ia, prim);
  else
    p2 ? launch_irs_xyz_p<true, true, TY>(c, L, ia, prim) : launch_irs_xyz_p<true, false, TY>(c, L, ia, prim);
}

template <typename R>
void launch_irs_xyz(fuss_ctx* c, Level& L, const IrsArgs<R>& ia, bool prim) {
  if (L.xyz_ty == 8) launch_irs_xyz_ty<R, 8>(c, L, ia, prim);
  else launch_irs_xyz_ty<R, 4>(c, L, ia, prim);
}

template <int DIR, int NDIR, typename R>
void launch_irs_d(fuss_ctx* c, Level& L, const IrsArgs<R>& a, bool update, bool prim) {
  const unsigned g = (unsigned)L.htiles.size();
  if (!update) k_irs<DIR, NDIR, false, true, R><<<g, TILE_T, 0, c->stream>>>(a);
  else if (prim) k_irs<DIR, NDIR, true, true, R><<<g, TILE_T, 0, c->stream>>>(a);
  else k_irs<DIR, NDIR, true, false, R><<<g, TILE_T, 0, c->stream>>>(a);
}

template <typename R>
void launch_irs_line(fuss_ctx* c, Level& L, int ndir, int dir, const IrsArgs<R>& a, bool update, bool prim) {
  if (ndir == 2) {
    if (dir == 0) launch_irs_d<0, 2>(c, L, a, update, prim);
    else launch_irs_d<1, 2>(c, L, a, update, prim);
  } else {
    if (dir == 0) launch_irs_d<0, 3>(c, L, a, update, prim);
    else if (dir == 1) launch_irs_d<1, 3>(c, L, a, update, prim);
    else launch_irs_d<2, 3>(c, L, a, update, prim);
  }
}

// x and y in one LDS pass; 2-D levels apply the update there
template <typename R>
void launch_irs_xy(fuss_ctx* c, Level& L, int ndir, const IrsArgs<R>& ia, bool prim) {
  const unsigned gxy = (unsigned)L.hxytiles.size();
  if (ndir == 2)
    prim ? k_irs_xy<2, true, true, R><<<gxy, TILE_T, 0, c->stream>>>(ia)
         : k_irs_xy<2, true, false, R><<<gxy, TILE_T, 0, c->stream>>>(ia);
  else
    prim ? k_irs_xy<3, false, true, R><<<gxy, TILE_T, 0, c->stream>>>(ia)
         : k_irs_xy<3, false, false, R><<<gxy, TILE_T, 0, c->stream>>>(ia);
}

// Residual_Smoothing with the tridiagonal lines: x, y (, z) passes
// rin -> rout -> rin (-> rout), each with the ghosts of the current R, then
// the RK update.  Sets the tile count of the update's residual partials.
template <typename R>
int launch_irs_tridiag(fuss_ctx* c, Level& L, IrsArgs<R> ia, R* rin, R* rout, bool prim, int* npart) {
  TriArgs<R> ta{};
  ta.blk = L.blk.p; ta.bc = L.bc.p; ta.fmap = L.fmap.p; ta.tri = L.tri.p; ta.trioff = L.trioff.p;
  ta.beta = ia.beta; ta.halt = ia.halt;
  for (int d = 0; d < c->ndir; d++) {
    ta.Rin = rin; ta.Rout = rout;
    const int pid = prof_begin(c, FUSS_K_IRS);
    if (d == 0) {
      ta.tiles = L.ttx.p;
      k_tri_x<R><<<(unsigned)L.htx.size(), 64, 0, c->stream>>>(ta);
    } else if (d == 1) {
      ta.tiles = L.tty.p;
      k_tri_yz<1, R><<<(unsigned)L.hty.size(), TILE_T, 0, c->stream>>>(ta);
    } else {
      ta.tiles = L.ttz.p;
      k_tri_yz<2, R><<<(unsigned)L.htz.size(), TILE_T, 0, c->stream>>>(ta);
    }
    prof_end(c, pid);
    HIPCHK(hipGetLastError());
    std::swap(rin, rout);
  }
  ia.Rin = rin;
  ia.tiles = L.tiles.p;
  const int pid = prof_begin(c, FUSS_K_IRS_UPDATE);
  prim ? k_rk_update<true, R><<<(unsigned)L.htiles.size(), TILE_T, 0, c->stream>>>(ia)
       : k_rk_update<false, R><<<(unsigned)L.htiles.size(), TILE_T, 0, c->stream>>>(ia);
  prof_end(c, pid);
  HIPCHK(hipGetLastError());
  *npart = (int)L.htiles.size();
  return 0;
}

}  // namespace fuss
