// fuss_stage_reg.h -- register-column stage kernel for rectilinear 3-D
// blocks (gfx950).  Included by the stage_reg_*.hip units.
//
// One wave owns a column of 64 x 4 cells (the level's 4-row column tiles)
// over KCT consecutive k-planes and marches up in k with no LDS and no
// barrier: lane l holds cell i0 + l of the column's four rows in registers.
//   x neighbours: DPP lane shifts (wave_shl / wave_shr); the right neighbour
//     of lane 63 comes from a per-row halo load, the x face on the column's
//     min side (lane 0) from one face evaluation per plane in which edge lane
//     e < 4 evaluates row e's face; a quad_perm move then brings row r's
//     value to lane 0 before the wave_shr of the +x fluxes.
//   y neighbours: the lane's own registers; rows j0-1 and j0+4 are loaded as
//     a six-row window, so 5 y faces serve 4 rows.
//   z neighbours: windows of planes k (cells), k+1 (z faces) rotate through
//     three register sets, the third receiving plane k+2; the +z flux of a
//     plane is the next plane's -z flux.
// Each face flux is evaluated once per column except the two edge faces of
// the column (3.5 evaluations per cell).  The arithmetic per face and cell,
// the accumulation order of the residual (x, y, z faces, then the BC values
// in face order) and the residual scaling are those of k_stage3p<RECT>
// (Mod_Fluxes.f90:54-124, Lib_Diffusive.f90:10-45, Lib_Newstate.f90:241-256),
// so the two kernels produce the same bits.
//
// Boundary faces.  A block face is never added: the y and z tests are
// wave-uniform branches; on x the flux of a face that is not interior is
// made exactly +-0 (zero weight), and adding +-0 to the running sum, which
// starts at +0 and can never become -0, leaves it unchanged.
#pragma once

#include "fuss_stage.h"

namespace fuss {

constexpr int RC_ROWS = TILE_Y;  // rows of a register column (the 4-row column tiles)
constexpr int RC_WAVES = 4;      // columns (waves) per work-group, independent of each other

// DPP moves of a value of type R across the wave.  old: the value kept by
// lanes whose source lane lies outside the wave (bound_ctrl off).
constexpr int DPP_WAVE_SHL1 = 0x130;  // lane l <- lane l+1
constexpr int DPP_WAVE_SHR1 = 0x138;  // lane l <- lane l-1
template <int CTRL>
__device__ __forceinline__ double dpp_mov(double old, double src) {
  const unsigned long long o = (unsigned long long)__double_as_longlong(old);
  const unsigned long long s = (unsigned long long)__double_as_longlong(src);
  const unsigned lo = (unsigned)__builtin_amdgcn_update_dpp((int)(unsigned)o, (int)(unsigned)s, CTRL, 0xf, 0xf, false);
  const unsigned hi =
      (unsigned)__builtin_amdgcn_update_dpp((int)(unsigned)(o >> 32), (int)(unsigned)(s >> 32), CTRL, 0xf, 0xf, false);
  return __longlong_as_double((long long)(((unsigned long long)hi << 32) | lo));
}
template <int CTRL>
__device__ __forceinline__ float dpp_mov(float old, float src) {
  return __int_as_float(__builtin_amdgcn_update_dpp(__float_as_int(old), __float_as_int(src), CTRL, 0xf, 0xf, false));
}
// lane 0 <- lane e (e < 4) of each quad (quad_perm [e,1,2,3]: every lane
// has a source, so no old value); e is a constant after unrolling
template <int CTRL>
__device__ __forceinline__ double dpp_get(double src) {
  const unsigned long long s = (unsigned long long)__double_as_longlong(src);
  const unsigned lo = (unsigned)__builtin_amdgcn_mov_dpp((int)(unsigned)s, CTRL, 0xf, 0xf, false);
  const unsigned hi = (unsigned)__builtin_amdgcn_mov_dpp((int)(unsigned)(s >> 32), CTRL, 0xf, 0xf, false);
  return __longlong_as_double((long long)(((unsigned long long)hi << 32) | lo));
}
template <int CTRL>
__device__ __forceinline__ float dpp_get(float src) {
  return __int_as_float(__builtin_amdgcn_mov_dpp(__float_as_int(src), CTRL, 0xf, 0xf, false));
}
template <typename R>
__device__ __forceinline__ R dpp_quad_from(R src, int e) {
  switch (e) {
    case 1: return dpp_get<0xE4 | 1>(src);
    case 2: return dpp_get<0xE4 | 2>(src);
    case 3: return dpp_get<0xE4 | 3>(src);
    default: return src;
  }
}

// Raw buffer accesses: a resource per array (base address, no range
// limit), a loop-invariant 32-bit lane offset in a VGPR and the plane's byte
// offset in an SGPR, so no 64-bit address is formed per row and array.
// aux 2: non-temporal (streams read or written once).
using Rsrc = __amdgpu_buffer_rsrc_t;
__device__ __forceinline__ Rsrc rsrc(const void* p) {
  return __builtin_amdgcn_make_buffer_rsrc(const_cast<void*>(p), (short)0, -1, 0x00020000);
}
typedef unsigned u32x2 __attribute__((__vector_size__(2 * sizeof(unsigned))));
template <typename T, int AUX = 0>
__device__ __forceinline__ T bld(Rsrc r, unsigned voff, unsigned soff) {
  if constexpr (sizeof(T) == 8) {
    return __builtin_bit_cast(T, __builtin_amdgcn_raw_buffer_load_b64(r, voff, soff, AUX));
  } else {
    return __builtin_bit_cast(T, __builtin_amdgcn_raw_buffer_load_b32(r, voff, soff, AUX));
  }
}
template <typename T, int AUX = 0>
__device__ __forceinline__ void bst(T v, Rsrc r, unsigned voff, unsigned soff) {
  if constexpr (sizeof(T) == 8) {
    __builtin_amdgcn_raw_buffer_store_b64(__builtin_bit_cast(u32x2, v), r, voff, soff, AUX);
  } else {
    __builtin_amdgcn_raw_buffer_store_b32(__builtin_bit_cast(unsigned, v), r, voff, soff, AUX);
  }
}

template <typename R>
__device__ __forceinline__ R ldbn(const void* base, unsigned off) {
  return __builtin_nontemporal_load(reinterpret_cast<const R*>(reinterpret_cast<const char*>(base) + off));
}
template <typename R>
__device__ __forceinline__ void stbn(void* base, unsigned off, R v) {
  __builtin_nontemporal_store(v, reinterpret_cast<R*>(reinterpret_cast<char*>(base) + off));
}

// A face temperature's kappa row request: fraction T - row and the row.
template <typename R>
struct KReq {
  typename Vec2<R>::T k;
  R f;
};

// Six rows (j0-1 .. j0+4) of one plane of T at the lane's i.
template <typename R>
struct RcWin {
  R r[RC_ROWS + 2];
};

// The x halos of a plane's four rows (lane 63: its right neighbour), the
// edge lanes' two cells and the materials of rows j0-1 .. j0+3: what the
// plane's table-row requests depend on, loaded one plane ahead.
template <typename R>
struct RcHalo {
  R h[RC_ROWS];
  R el, er;
  int m[RC_ROWS + 1];
  int me;
};

__device__ __forceinline__ void sched_fence() { __builtin_amdgcn_sched_barrier(0); }

// UM: every block of the level has one material; QU: every block has a
// uniform qvol (one cached value instead of a stream); DTG: the global dt of
// a time-accurate step (no dt stream); DTC: the first stage of a steady
// primitive step computes the local dt (Compute_dt, Mod_dt.f90:26-84).
//
// Schedule of plane k: (A) every table-row request of the plane -- its y, x
// and edge faces (plane k and its x halos), its z faces (planes k and k+1)
// and its cells -- in one batch, so a plane waits for one round trip to the
// caches; (B) the plane's dt / T^n / qvol, then the window of plane k+2 and
// the halos of plane k+1, both used one plane later; (C) fluxes, residuals
// and stores.  The windows rotate through three register sets, the halos
// through two (six planes per trip keep every role static).
template <bool UM, bool QU, bool PRIM, bool IRSMODE, bool DTC, bool DTG, typename R, int KCT>
__global__ __launch_bounds__(64 * RC_WAVES) __attribute__((amdgpu_waves_per_eu(2)))
void k_stage3r(StageArgs<R> a, int ntile) {
#pragma clang fp contract(off)
  if (a.halt && *a.halt) return;
  constexpr int NR = RC_ROWS;
  constexpr unsigned SR = sizeof(R);
  const int lane = threadIdx.x & 63;
  const int wv = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int tid = xcd_remap(blockIdx.x, gridDim.x) * RC_WAVES + wv;
  if (tid >= ntile) return;
  const Tile t = a.ctiles[tid];
  const BlockInfo B = a.blk[t.b];
  const int ni = B.ni, nj = B.nj, nk = B.nk;
  const int i0 = t.i0, j0 = t.j0, k0 = t.k0, k1 = min(nk, k0 + KCT - 1);
  const int i = i0 + lane;
  const bool li = i <= ni;
  const int sy = B.sy;
  const long sz = B.sz;
  using V2 = typename Vec2<R>::T;
  using V4 = typename Vec4<R>::T;
  const int Ti1 = a.tab.Ti1, Ti2 = a.tab.Ti2, nT = a.tab.nT;
  const int Ti2m1 = to_vgpr(a.tab.Ti2 - 1);
  const int mu = B.umat >= 0 ? B.umat : 0;
  // kappa and (rho, drho, cp, dcp) rows addressed by the row number itself
  const V2* kapb = reinterpret_cast<const V2*>(a.tab.kap2) - Ti1;
  const V4* prb = reinterpret_cast<const V4*>(a.tab.pr4) - Ti1;
  const V2* kapu = kapb + (long)mu * nT;
  const V4* pru = prb + (long)mu * nT;
  // Plane origin: cell (1, j0-1, k).  Byte offsets from it of the lane's
  // cell in the window rows j0-1 .. j0+4 (rows clamped to the ghost row
  // nj+1, lanes to the ghost column ni+1); constant over the column.
  auto org = [&](int kk) -> long { return B.base + (long)(j0 - 2) * sy + (long)(kk - 1) * sz; };
  const int ec = min(i, ni + 1) - 1;
  unsigned ro[NR + 2];
#pragma unroll
  for (int r = 0; r < NR + 2; r++) ro[r] = (unsigned)(ec + (min(j0 - 1 + r, nj + 1) - (j0 - 1)) * sy) * SR;
  // x halo of a row, relative to the lane's cell: lane 63 reads i0+64 (its
  // right neighbour), the other lanes cell i0 (a line they read anyway)
  const int hc = lane == 63 ? min(i0 + 64, ni + 1) - 1 : i0 - 1;
  const unsigned hd = (unsigned)(hc - ec) * SR;
  // edge lane e < 4: cell (i0-1, j0+e) (the ghost i = 0 when i0 = 1, whose
  // face then has weight 0); its +1 neighbour is (i0, j0+e)
  const int er = min(lane, NR - 1);
  const unsigned eo = (unsigned)((min(j0 + er, nj + 1) - (j0 - 1)) * sy + i0 - 2) * SR;
  const bool edge_live = lane < NR && j0 + lane <= nj && i0 > 1;
  // Per-axis metric factors.  The lane's plane-independent products live in
  // LDS (one slot array per wave, read back per plane), not in registers:
  // cv[r] = a_x a_y (volume, z weight), c0[r] = f_x a_y (+x weight), c4[q] =
  // f_y a_x (+y weight of row j0+q-1), ce = f_x(i0-1) a_y (the edge lane's
  // face), qxy[r] = 1/dl_x^2 + 1/dl_y^2 (dt).  A face that is not interior
  // (the block's faces, rows and columns beyond it) gets weight 0.
  constexpr int SL_CV = 0, SL_C0 = NR, SL_C4 = 2 * NR, SL_CE = 3 * NR + 1, SL_QXY = 3 * NR + 2;
  constexpr int NSLOT = DTC ? 4 * NR + 2 : 3 * NR + 2;
  __shared__ double slots[RC_WAVES][NSLOT][64];
  double(*sl)[64] = slots[wv];
  const RectAxes rx = rect_axes(a.rect, B);
  {
    const int ci = li ? i : ni;  // (dead lanes read a valid factor)
    const double axi = rx.ax[ci];
    const double fxi = i < ni ? rx.fx[ci] : 0.0;
#pragma unroll
    for (int r = 0; r < NR; r++) {
      const double ayj = rx.ay[min(j0 + r, nj + 1)];
      sl[SL_CV + r][lane] = (double)(R)(axi * ayj);
      sl[SL_C0 + r][lane] = (double)(R)(fxi * ayj);
      if (DTC) sl[SL_QXY + r][lane] = rx.qx[ci] + rx.qy[min(j0 + r, nj + 1)];
    }
#pragma unroll
    for (int q = 0; q <= NR; q++) {
      const int jl = j0 - 1 + q;  // the face's left row
      sl[SL_C4 + q][lane] = jl >= 1 && jl < nj ? (double)(R)(rx.fy[jl] * axi) : 0.0;
    }
    const double fxe = i0 > 1 ? rx.fx[i0 - 1] : 0.0;
    sl[SL_CE][lane] = (double)(R)(fxe * rx.ay[min(j0 + er, nj + 1)]);
  }
  const double* const zax = rx.az;
  const double* const zfx = rx.fz;
  const double* const zqx = rx.qz;
  // BC values at the lane's boundary faces (face-map positions; at most one
  // face per axis since every block dimension is >= 2)
  const int fa = B.face_active;
  const int xsel = li ? ((fa & 1) && i == 1 ? B.fmap_off[0] : ((fa & 2) && i == ni ? B.fmap_off[1] : -1)) : -1;
  const bool xtile = (fa & 3) && (i0 == 1 || i0 + 63 >= ni);
  // x tiles: the lane's BC value at xb + xm * ((j-1) + nj (k-1)) (lanes
  // without one read the entry that stays 0)
  const unsigned xb = xsel >= 0 ? (unsigned)xsel : a.bcz;
  const unsigned xm = xsel >= 0 ? 1u : 0u;
  const R dtg = DTG ? (R)(*a.dtg) : (R)0;
  const R quni = QU ? a.qvol[B.base] : (R)0;
  const Rsrc rT = rsrc(a.T), rD = rsrc(a.dt), rTO = rsrc(a.TO), rQ = rsrc(a.qvol), rM = rsrc(a.mat);
  const Rsrc rOut = rsrc(IRSMODE ? a.Rres : a.Tout), rDw = rsrc(a.dtw);
  // a cell array's byte offset of plane kk's origin (int arrays: mo)
  auto so = [&](int kk) -> unsigned { return (unsigned)org(kk) * SR; };
  auto mo = [&](int kk) -> unsigned { return (unsigned)org(kk) * (unsigned)sizeof(int); };
  constexpr unsigned MS = SR / sizeof(int);  // lane offset scale of int arrays

  // kappa row request of a face temperature (left cell material m)
  auto kreq = [&](R Tf, int m) -> KReq<R> {
    const int row = row_sat(Tf, Ti1, Ti2m1);
    KReq<R> q;
    q.k = ldu(UM ? kapu : kapb + (long)m * nT, (unsigned)row);
    q.f = frac_of(Tf);  // Tf - row (an error is flagged where they differ)
    return q;
  };
  auto kin = [&](R T) -> int { return !(T >= (R)Ti1 && T < (R)Ti2); };
  auto kval = [&](const KReq<R>& q) -> R { return q.k.x + q.f * q.k.y; };
  auto load_win = [&](RcWin<R>& W, int kk, int r0, int r1) {
    const unsigned o = so(kk);
#pragma unroll
    for (int r = 0; r < NR + 2; r++)
      if (r >= r0 && r <= r1) W.r[r] = bld<R>(rT, ro[r], o);
  };
  auto load_halo = [&](RcHalo<R>& H, int kk) {
    const unsigned o = so(kk);
#pragma unroll
    for (int r = 0; r < NR; r++) H.h[r] = bld<R>(rT, ro[r + 1] + hd, o);
    H.el = bld<R>(rT, eo, o);
    H.er = bld<R>(rT, eo + SR, o);
    if (!UM) {
      const unsigned om = mo(kk);
#pragma unroll
      for (int r = 0; r <= NR; r++) H.m[r] = bld<int>(rM, ro[r] / MS, om);
      H.me = bld<int>(rM, eo / MS, om);
    }
  };

  RcWin<R> W0, W1, W2;
  RcHalo<R> H0, H1;
#pragma unroll
  for (int r = 0; r <= NR; r++) H0.m[r] = H1.m[r] = mu;
  H0.me = H1.me = mu;
  // -z fluxes of the column's first plane (+0 at the block's first plane)
  R fzp[NR];
#pragma unroll
  for (int r = 0; r < NR; r++) fzp[r] = (R)0;
  if (k0 > 1) {
    // W2 holds plane k0-1 for this prologue only (rows j0 .. j0+3)
    load_win(W2, k0 - 1, 1, NR);
    load_win(W0, k0, 0, NR + 1);
    const unsigned om = mo(k0 - 1);
    const double fzk0 = zfx[k0 - 1];
#pragma unroll
    for (int r = 0; r < NR; r++) {
      const R Tl = W2.r[r + 1], Tu = W0.r[r + 1];
      const int m = UM ? mu : bld<int>(rM, ro[r + 1] / MS, om);
      const KReq<R> q = kreq((R)0.5 * (Tl + Tu), m);
      const R w8 = (R)(fzk0 * sl[SL_CV + r][lane]);
      fzp[r] = kval(q) * ((Tu - Tl) * w8);
    }
  } else {
    load_win(W0, k0, 0, NR + 1);
  }
  load_halo(H0, k0);
  load_win(W1, k0 + 1, 0, NR + 1);

  double rsq = 0.0, dtmin = 1e300;
  int ebad = 0;  // first cell of the lane with a table error: (k << 3) | 4 | row

  // One plane: Wc = plane k, Wn = plane k+1, Hc = plane k's halos (all
  // requested a plane earlier); Wf receives plane k+2, Hn the halos of plane
  // k+1.  The column's last plane requests valid planes it then ignores.
  auto plane = [&](RcWin<R>& Wc, RcWin<R>& Wn, RcWin<R>& Wf, RcHalo<R>& Hc, RcHalo<R>& Hn, int k) {
    const double azk = zax[k];
    const double fzk = k < nk ? zfx[k] : 0.0;  // (0: the block's last plane)
    int badf = 0;
    // ---- (A) table-row requests of the plane
    KReq<R> qy[NR + 1], qx[NR], qe, qz[NR];
    V4 pc[NR];
    V2 kc[NR];
    R Tr[NR];
#pragma unroll
    for (int q = 0; q <= NR; q++) {  // y face q: rows j0+q-1 | j0+q
      const R Tf = (R)0.5 * (Wc.r[q] + Wc.r[q + 1]);
      qy[q] = kreq(Tf, Hc.m[q]);
      if (!PRIM) badf |= li && j0 + q - 1 >= 1 && j0 + q <= nj && kin(Tf);
    }
#pragma unroll
    for (int r = 0; r < NR; r++) {
      const R Tc = Wc.r[r + 1];
      Tr[r] = dpp_mov<DPP_WAVE_SHL1>(Hc.h[r], Tc);
      const R Tf = (R)0.5 * (Tc + Tr[r]);
      qx[r] = kreq(Tf, Hc.m[r + 1]);
      if (!PRIM) badf |= li && i < ni && j0 + r <= nj && kin(Tf);
    }
    {
      const R Tf = (R)0.5 * (Hc.el + Hc.er);
      qe = kreq(Tf, Hc.me);
      if (!PRIM) badf |= edge_live && kin(Tf);
    }
#pragma unroll
    for (int r = 0; r < NR; r++) {
      const R Tf = (R)0.5 * (Wc.r[r + 1] + Wn.r[r + 1]);
      qz[r] = kreq(Tf, Hc.m[r + 1]);
      if (!PRIM) badf |= li && k < nk && j0 + r <= nj && kin(Tf);
    }
    int badc = 0;
    if (PRIM) {
#pragma unroll
      for (int r = 0; r < NR; r++) {
        const R Tc = Wc.r[r + 1];
        const int row = row_sat(Tc, Ti1, Ti2m1);
        pc[r] = ldu(UM ? pru : prb + (long)Hc.m[r + 1] * nT, (unsigned)row);
        if (DTC) kc[r] = ldu(UM ? kapu : kapb + (long)Hc.m[r + 1] * nT, (unsigned)row);
        badc |= (li && j0 + r <= nj && kin(Tc)) << r;
      }
    }
    if (!PRIM && badf) flag_error(a.status, FUSS_ETABLE, -1, 0, 0, 0);
    // ---- (B) the plane's streams, then the next planes' window and halos
    R dtc[NR], toc[NR], qvc[NR];
    {
      const unsigned o = so(k);
#pragma unroll
      for (int r = 0; r < NR; r++) {
        dtc[r] = DTG || DTC ? dtg : bld<R, 2>(rD, ro[r + 1], o);
        toc[r] = IRSMODE ? (R)0 : bld<R, 2>(rTO, ro[r + 1], o);
        qvc[r] = QU ? quni : bld<R, 2>(rQ, ro[r + 1], o);
      }
    }
    load_win(Wf, min(k + 2, nk + 1), 0, NR + 1);
    load_halo(Hn, min(k + 1, nk));
    sched_fence();
    // ---- (C) fluxes, residuals, updates
    R fy[NR + 1];
#pragma unroll
    for (int q = 0; q <= NR; q++) {
      const R w4 = (R)(sl[SL_C4 + q][lane] * azk);
      fy[q] = kval(qy[q]) * ((Wc.r[q + 1] - Wc.r[q]) * w4);
    }
    const R we = (R)(sl[SL_CE][lane] * azk);
    const R fe = kval(qe) * ((Hc.er - Hc.el) * we);
    const int zsel = (fa & 16) && k == 1 ? B.fmap_off[4] : ((fa & 32) && k == nk ? B.fmap_off[5] : -1);
    const int il = li ? i - 1 : 0;
    R res[NR], dtr[NR];
#pragma unroll
    for (int r = 0; r < NR; r++) {
      const int j = j0 + r;
      res[r] = dtr[r] = (R)0;
      const R Tc = Wc.r[r + 1];
      const R w0 = (R)(sl[SL_C0 + r][lane] * azk);
      const R fxp = kval(qx[r]) * ((Tr[r] - Tc) * w0);
      // -x flux: the +x flux of lane l-1; lane 0 takes edge lane r's value
      const R fxm = dpp_mov<DPP_WAVE_SHR1>(dpp_quad_from(fe, r), fxp);
      if (j > nj) continue;
      const double cvol = sl[SL_CV + r][lane];
      const R w8 = (R)(fzk * cvol);
      const R fz = kval(qz[r]) * ((Wn.r[r + 1] - Tc) * w8);
      R v = ((((fxm - fxp) + fy[r]) - fy[r + 1]) + fzp[r]) - fz;
      fzp[r] = fz;
      const int ysel = (fa & 4) && j == 1 ? B.fmap_off[2] : ((fa & 8) && j == nj ? B.fmap_off[3] : -1);
      if (xtile || ysel >= 0 || zsel >= 0) {
        // BC_Fluxes values in face order (as apply_bc_faces); +0 where the
        // cell has none
        R bx = 0, by = 0, bz = 0;
        if (xtile) bx = ldu(a.bcf, xb + xm * (unsigned)((j - 1) + nj * (k - 1)));
        if (ysel >= 0) by = ldu(a.bcf, (unsigned)(ysel + il + ni * (k - 1)));
        if (zsel >= 0) bz = ldu(a.bcf, (unsigned)(zsel + il + ni * (j - 1)));
        v = v - bx;
        v = v - by;
        v = v - bz;
      }
      const R vol = (R)(cvol * azk);
      R dt = dtc[r];
      R fc = 0;
      if (PRIM) fc = frac_of(Tc);
      if (DTC) {
        const R kap = kc[r].x + fc * kc[r].y, rho = pc[r].x + fc * pc[r].y, cp = pc[r].z + fc * pc[r].w;
        const double alpha = (double)kap / (double)rho / (double)cp;
        double dd = 0.5 * a.vnn / (alpha * (sl[SL_QXY + r][lane] + zqx[k]));
        dd = dd < 1e8 ? dd : 1e8;
        dt = (R)dd;
        dtr[r] = dt;
        if (li) dtmin = fmin(dtmin, dd);
      }
      R rs;
      if (PRIM) {
        const R rho = pc[r].x + fc * pc[r].y;
        const R cp = pc[r].z + fc * pc[r].w;
        rs = (-v / vol + qvc[r]) * dt / (Tc * cp * pc[r].y + rho * cp + Tc * rho * pc[r].w);
      } else {
        rs = (-v / vol + qvc[r]) * dt;
      }
      int bad = (badc >> r) & 1;
      if (IRSMODE) {
        res[r] = rs;
      } else {
        int bad2 = 0;
        const R TO = toc[r];
        const R Tn = new_state<PRIM>(a.tab, Hc.m[r + 1], Tc, TO, rs, a.rkc, a.jam, &bad2);
        res[r] = Tn;
        if (li) {
          if (Tn != Tn || Tn < (R)0) flag_state(a.status, a.irk, t.b, i, j, k);
          const double d = (double)Tn - (double)TO;
          rsq += d * d;
        }
        bad |= bad2;
      }
      // the lane's first table error, reported at the column's end
      if (li && bad && !ebad) ebad = (k << 3) | 4 | r;
    }
    if (li) {
      const unsigned o = so(k);
#pragma unroll
      for (int r = 0; r < NR; r++) {
        if (j0 + r > nj) continue;
        bst<R, 2>(res[r], rOut, ro[r + 1], o);
        if (DTC) bst<R, 2>(dtr[r], rDw, ro[r + 1], o);
      }
    }
    sched_fence();
  };
  for (int k = k0; k <= k1; k += 6) {
    plane(W0, W1, W2, H0, H1, k);
    if (k + 1 > k1) break;
    plane(W1, W2, W0, H1, H0, k + 1);
    if (k + 2 > k1) break;
    plane(W2, W0, W1, H0, H1, k + 2);
    if (k + 3 > k1) break;
    plane(W0, W1, W2, H1, H0, k + 3);
    if (k + 4 > k1) break;
    plane(W1, W2, W0, H0, H1, k + 4);
    if (k + 5 > k1) break;
    plane(W2, W0, W1, H1, H0, k + 5);
  }
  if (ebad) flag_error(a.status, FUSS_ETABLE, t.b, i, j0 + (ebad & 3), ebad >> 3);
  if (!IRSMODE && a.last) {
    for (int o = 32; o > 0; o >>= 1) rsq += __shfl_xor(rsq, o, 64);
    if (lane == 0) a.partials[tid] = rsq;
  }
  if (DTC) {
    for (int o = 32; o > 0; o >>= 1) dtmin = fmin(dtmin, __shfl_xor(dtmin, o, 64));
    if (lane == 0) a.pmin[tid] = dtmin;
  }
}

}  // namespace fuss
