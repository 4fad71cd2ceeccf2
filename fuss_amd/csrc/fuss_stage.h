// fuss_stage.h -- stage kernels (interior fluxes + BC faces + newstate) of
// the FUSS explicit step on gfx950.  Included by the stage launch units.
#pragma once

#include <type_traits>

#include "fuss_dev.h"

template <int NDIR, bool TANG, bool PRIM, bool IRSMODE, typename R>
__global__ __launch_bounds__(TILE_T) void k_stage(StageArgs<R> a) {
  if (a.halt && *a.halt) return;
  const int wg = xcd_remap(blockIdx.x, gridDim.x);
  const Tile t = a.tiles[wg];
  const BlockInfo B = a.blk[t.b];
  const int i = t.i0 + (threadIdx.x % TILE_X), j = t.j0 + (threadIdx.x / TILE_X), k = t.k0;
  double rsq = 0.0;
  if (i <= B.ni && j <= B.nj) {
    const long p = B.base + (i - 1) + (long)(j - 1) * B.sy + (long)(k - 1) * B.sz;
    R Rv = 0;
    {
#pragma clang fp contract(off)
      // Fluxes_blk ordering (Mod_Fluxes.f90:54-192): i faces, then j, then k;
      // Res2 += F for the face on the min side, Res1 -= F on the max side.
      if (i > 1) Rv = Rv + face_flux<0, NDIR, TANG>(a, B, p - 1);
      if (i < B.ni) Rv = Rv - face_flux<0, NDIR, TANG>(a, B, p);
      if (j > 1) Rv = Rv + face_flux<1, NDIR, TANG>(a, B, p - B.sy);
      if (j < B.nj) Rv = Rv - face_flux<1, NDIR, TANG>(a, B, p);
      if (NDIR == 3) {
        if (k > 1) Rv = Rv + face_flux<2, NDIR, TANG>(a, B, p - B.sz);
        if (k < B.nk) Rv = Rv - face_flux<2, NDIR, TANG>(a, B, p);
      }
    }
    Rv = apply_bc_faces(a, B, i, j, k, Rv);
    const R T = a.T[p];
    const int m = a.mat[p];
    const R dt = a.use_dtg ? (R)(*a.dtg) : a.dt[p];
    int bad = 0;
    const R res = scaled_residual<PRIM>(a.tab, m, T, Rv, a.vol[p], a.qvol[p], dt, &bad);
    if (IRSMODE) {
      a.Rres[p] = res;
    } else {
      const R TO = a.TO[p];
      const R Tn = new_state<PRIM>(a.tab, m, T, TO, res, a.rkc, a.jam, &bad);
      a.Tout[p] = Tn;
      // Check_And_Fix_State, Lib_Newstate.f90:278-290
      if (Tn != Tn || Tn < (R)0) flag_state(a.status, a.irk, t.b, i, j, k);
      const double d = (double)Tn - (double)TO;
      rsq = d * d;
    }
    if (bad) flag_error(a.status, FUSS_ETABLE, t.b, i, j, k);
  }
  if (!IRSMODE && a.last) block_sum(rsq, a.partials, wg);
}

// The first half of k_stage for a step with a host hook between BC_Fluxes
// and RK_Newstate (External_Function, Mod_Explicit.f90:53-57): the raw
// residual R of Fluxes + BC_Fluxes, in k_stage's order, stored to Rres.
template <int NDIR, bool TANG, typename R>
__global__ __launch_bounds__(TILE_T) void k_stage_raw(StageArgs<R> a) {
  if (a.halt && *a.halt) return;
  const Tile t = a.tiles[blockIdx.x];
  const BlockInfo B = a.blk[t.b];
  const int i = t.i0 + (threadIdx.x % TILE_X), j = t.j0 + (threadIdx.x / TILE_X), k = t.k0;
  if (i > B.ni || j > B.nj) return;
  const long p = B.base + (i - 1) + (long)(j - 1) * B.sy + (long)(k - 1) * B.sz;
  R Rv = 0;
  {
#pragma clang fp contract(off)
    if (i > 1) Rv = Rv + face_flux<0, NDIR, TANG>(a, B, p - 1);
    if (i < B.ni) Rv = Rv - face_flux<0, NDIR, TANG>(a, B, p);
    if (j > 1) Rv = Rv + face_flux<1, NDIR, TANG>(a, B, p - B.sy);
    if (j < B.nj) Rv = Rv - face_flux<1, NDIR, TANG>(a, B, p);
    if (NDIR == 3) {
      if (k > 1) Rv = Rv + face_flux<2, NDIR, TANG>(a, B, p - B.sz);
      if (k < B.nk) Rv = Rv - face_flux<2, NDIR, TANG>(a, B, p);
    }
  }
  a.Rres[p] = apply_bc_faces(a, B, i, j, k, Rv);
}

// The second half: RK_Newstate from the (hook-modified) R in Rres, as k_stage
// continues (IRSMODE: the scaled residual replaces R for the smoothing passes).
template <bool PRIM, bool IRSMODE, typename R>
__global__ __launch_bounds__(TILE_T) void k_newstate(StageArgs<R> a) {
  if (a.halt && *a.halt) return;
  const int wg = blockIdx.x;
  const Tile t = a.tiles[wg];
  const BlockInfo B = a.blk[t.b];
  const int i = t.i0 + (threadIdx.x % TILE_X), j = t.j0 + (threadIdx.x / TILE_X), k = t.k0;
  double rsq = 0.0;
  if (i <= B.ni && j <= B.nj) {
    const long p = B.base + (i - 1) + (long)(j - 1) * B.sy + (long)(k - 1) * B.sz;
    const R Rv = a.Rres[p];
    const R T = a.T[p];
    const int m = a.mat[p];
    const R dt = a.use_dtg ? (R)(*a.dtg) : a.dt[p];
    int bad = 0;
    const R res = scaled_residual<PRIM>(a.tab, m, T, Rv, a.vol[p], a.qvol[p], dt, &bad);
    if (IRSMODE) {
      a.Rres[p] = res;
    } else {
      const R TO = a.TO[p];
      const R Tn = new_state<PRIM>(a.tab, m, T, TO, res, a.rkc, a.jam, &bad);
      a.Tout[p] = Tn;
      if (Tn != Tn || Tn < (R)0) flag_state(a.status, a.irk, t.b, i, j, k);
      const double d = (double)Tn - (double)TO;
      rsq = d * d;
    }
    if (bad) flag_error(a.status, FUSS_ETABLE, t.b, i, j, k);
  }
  if (!IRSMODE && a.last) block_sum(rsq, a.partials, wg);
}


// ------------------------------------------------------------------------
// 2.5-D stage kernel for 3-D levels: a work-group owns a 64x4 column of
// cells over KC consecutive k-planes and marches up in k.  Three planes of T
// (k-1, k, k+1, each with a one-cell halo in i and j) rotate through LDS;
// every face flux is evaluated once: +x / +y fluxes go through LDS to the
// neighbouring cell, the +z flux is carried in a register to the next plane.
// The residual accumulation keeps the Fluxes_blk order (i, then j, then k
// faces; Mod_Fluxes.f90:54-192).
// ------------------------------------------------------------------------
constexpr int PX = TILE_X + 2, PY = TILE_Y + 2;

template <typename R>
__device__ __forceinline__ void load_plane(R (*dst)[PX], const R* __restrict__ T, const BlockInfo& B, int i0,
                                           int j0, int kk) {
  const long base = B.base + (long)(kk - 1) * B.sz;
  for (int q = threadIdx.x; q < PX * PY; q += TILE_T) {
    const int ii = q % PX, jj = q / PX;
    const int gi = i0 - 1 + ii, gj = j0 - 1 + jj;
    R v = 0;
    if (gi <= B.ni + 1 && gj <= B.nj + 1) v = T[base + (gi - 1) + (long)(gj - 1) * B.sy];
    dst[jj][ii] = v;
  }
}

// A plane with halo has PX*PY = 396 values: thread t holds elements t and
// t + TILE_T (when < 396) in registers between the global load and the LDS
// store, so the load of plane k+2 overlaps the flux work of plane k.
template <typename R>
struct PlaneRegs {
  R v0, v1;
};
template <typename R>
__device__ __forceinline__ void fetch_plane(PlaneRegs<R>& r, const R* __restrict__ T, const BlockInfo& B, int i0,
                                            int j0, int kk) {
  // arena index of the halo corner (i0-1, j0-1, kk); all offsets stay 32-bit
  const unsigned corner = (unsigned)(B.base + (i0 - 2) + (long)(j0 - 2) * B.sy + (long)(kk - 1) * B.sz);
  {
    const int q = threadIdx.x, ii = q % PX, jj = q / PX;
    r.v0 = (i0 - 1 + ii <= B.ni + 1 && j0 - 1 + jj <= B.nj + 1) ? ldu(T, corner + ii + jj * B.sy) : (R)0;
  }
  {
    const int q = threadIdx.x + TILE_T, ii = q % PX, jj = q / PX;
    r.v1 = (q < PX * PY && i0 - 1 + ii <= B.ni + 1 && j0 - 1 + jj <= B.nj + 1) ? ldu(T, corner + ii + jj * B.sy)
                                                                                : (R)0;
  }
}
// TY: rows of the tile (planes of PX x (TY+2) values; 64*TY threads, two
// values each)
template <int TY = TILE_Y, typename R>
__device__ __forceinline__ void store_plane(R (*dst)[PX], const PlaneRegs<R>& r) {
  constexpr int TT = TILE_X * TY;
  (&dst[0][0])[threadIdx.x] = r.v0;
  if (threadIdx.x + TT < PX * (TY + 2)) (&dst[0][0])[threadIdx.x + TT] = r.v1;
}

// kappa row of the left cell of a face: blocks of one material skip the
// per-cell material load
template <typename R>
__device__ __forceinline__ const R* kap_row(const StageArgs<R>& a, const BlockInfo& B, unsigned q) {
  const int m = B.umat >= 0 ? B.umat : ldu(a.mat, q);
  return a.tab.kap2 + 2 * (long)m * a.tab.nT;
}

// Face fluxes from the LDS planes (m = k-1, c = k, p = k+1); (x, y) are the
// LDS coordinates of the left cell, q its arena index.  Same algebra as
// face_flux (Lib_Diffusive.f90:10-45 regrouped per computational direction).
template <bool TANG, typename R>
__device__ __forceinline__ R flux_x(const StageArgs<R>& a, const BlockInfo& B, R (*m)[PX], R (*c)[PX], R (*p)[PX],
                                    int x, int y, unsigned q, R wn) {
  const R Tl = c[y][x], Tr = c[y][x + 1];
  int bad = 0;
  const R kap = tab_interp2(kap_row(a, B, q), (R)0.5 * (Tl + Tr), a.tab.Ti1, a.tab.Ti2, &bad);
  if (bad) flag_error(a.status, FUSS_ETABLE, -1, 0, 0, 0);
  const R G0 = Tr - Tl;
  if (!TANG) return kap * (G0 * wn);
  const R G1 = (c[y + 1][x] - c[y - 1][x] + c[y + 1][x + 1] - c[y - 1][x + 1]) * (R)0.25;
  const R G2 = (p[y][x] - m[y][x] + p[y][x + 1] - m[y][x + 1]) * (R)0.25;
  return kap * ((G0 * ldu(a.W[0], q) + G1 * ldu(a.W[1], q)) + G2 * ldu(a.W[2], q));
}

template <bool TANG, typename R>
__device__ __forceinline__ R flux_y(const StageArgs<R>& a, const BlockInfo& B, R (*m)[PX], R (*c)[PX], R (*p)[PX],
                                    int x, int y, unsigned q, R wn) {
  const R Tl = c[y][x], Tr = c[y + 1][x];
  int bad = 0;
  const R kap = tab_interp2(kap_row(a, B, q), (R)0.5 * (Tl + Tr), a.tab.Ti1, a.tab.Ti2, &bad);
  if (bad) flag_error(a.status, FUSS_ETABLE, -1, 0, 0, 0);
  const R G1 = Tr - Tl;
  if (!TANG) return kap * (G1 * wn);
  const R G0 = (c[y][x + 1] - c[y][x - 1] + c[y + 1][x + 1] - c[y + 1][x - 1]) * (R)0.25;
  const R G2 = (p[y][x] - m[y][x] + p[y + 1][x] - m[y + 1][x]) * (R)0.25;
  return kap * ((G0 * ldu(a.W[3], q) + G1 * ldu(a.W[4], q)) + G2 * ldu(a.W[5], q));
}

template <bool TANG, typename R>
__device__ __forceinline__ R flux_z(const StageArgs<R>& a, const BlockInfo& B, R (*c)[PX], R (*p)[PX], int x, int y,
                                    unsigned q, R wn) {
  const R Tl = c[y][x], Tr = p[y][x];
  int bad = 0;
  const R kap = tab_interp2(kap_row(a, B, q), (R)0.5 * (Tl + Tr), a.tab.Ti1, a.tab.Ti2, &bad);
  if (bad) flag_error(a.status, FUSS_ETABLE, -1, 0, 0, 0);
  const R G2 = Tr - Tl;
  if (!TANG) return kap * (G2 * wn);
  const R G0 = (c[y][x + 1] - c[y][x - 1] + p[y][x + 1] - p[y][x - 1]) * (R)0.25;
  const R G1 = (c[y + 1][x] - c[y - 1][x] + p[y + 1][x] - p[y - 1][x]) * (R)0.25;
  return kap * ((G0 * ldu(a.W[6], q) + G1 * ldu(a.W[7], q)) + G2 * ldu(a.W[8], q));
}

template <bool TANG, bool PRIM, bool IRSMODE, typename R, int KCT>
__global__ __launch_bounds__(TILE_T) void k_stage3d(StageArgs<R> a) {
  if (a.halt && *a.halt) return;
  const int wg = xcd_remap(blockIdx.x, gridDim.x);
  const Tile t = a.ctiles[wg];
  const BlockInfo B = a.blk[t.b];
  const int tx = threadIdx.x % TILE_X, ty = threadIdx.x / TILE_X;
  const int i = t.i0 + tx, j = t.j0 + ty;
  const int k0 = t.k0, k1 = min(B.nk, k0 + KCT - 1);
  const bool live = i <= B.ni && j <= B.nj;
  const int x = tx + 1, y = ty + 1;  // LDS coordinates of the thread's cell
  // min-side halo faces of the tile: wave 0 computes the y-halo (one face per
  // lane), wave 1 lanes 0..TILE_Y-1 the x-halo of rows 0..TILE_Y-1
  const bool yhalo = ty == 0 && t.j0 > 1 && i <= B.ni;
  const int hr = tx;  // x-halo row handled by this lane of wave 1
  const bool xhalo = ty == 1 && tx < TILE_Y && t.i0 > 1 && t.j0 + hr <= B.nj;
  __shared__ R ring[3][PY][PX];
  __shared__ R fx[TILE_Y][TILE_X + 1];
  __shared__ R fy[TILE_Y + 1][TILE_X];
  int sm = 0, sc = 1, sp = 2;
  PlaneRegs<R> pre;
  fetch_plane(pre, a.T, B, t.i0, t.j0, k0 - 1);
  store_plane(ring[sm], pre);
  fetch_plane(pre, a.T, B, t.i0, t.j0, k0);
  store_plane(ring[sc], pre);
  fetch_plane(pre, a.T, B, t.i0, t.j0, k0 + 1);
  __syncthreads();
  const unsigned sy = (unsigned)B.sy, sz = (unsigned)B.sz;
  unsigned p = (unsigned)(B.base + (i - 1) + (long)(j - 1) * B.sy + (long)(k0 - 1) * B.sz);
  // arena index of the x-halo cell (i0-1, j0+hr, k) of wave 1
  unsigned ph = (unsigned)(B.base + (t.i0 - 2) + (long)(t.j0 + hr - 1) * B.sy + (long)(k0 - 1) * B.sz);
  R fz_prev = 0;
  if (live && k0 > 1) {
#pragma clang fp contract(off)
    fz_prev = flux_z<TANG>(a, B, ring[sm], ring[sc], x, y, p - sz, TANG ? (R)0 : ldu(a.W[8], p - sz));
  }
  // normal face weights of the thread's cell, fetched one plane ahead
  // (orthogonal blocks read only these three weights)
  R w0 = 0, w4 = 0, w8 = 0;
  if (!TANG && live) {
    w0 = ldu(a.W[0], p);
    w4 = ldu(a.W[4], p);
    w8 = ldu(a.W[8], p);
  }
  double rsq = 0.0;
  for (int k = k0; k <= k1; k++, p += sz, ph += sz) {
    store_plane(ring[sp], pre);
    __syncthreads();
    if (k < k1) fetch_plane(pre, a.T, B, t.i0, t.j0, k + 2);
    // per-cell streams of this plane, issued before the flux work so that
    // their latency overlaps it
    R c_vol = 0, c_qv = 0, c_dt = 0, c_TO = 0;
    int c_m = B.umat;
    if (live) {
      c_vol = ldu(a.vol, p);
      c_qv = ldu(a.qvol, p);
      if (!a.use_dtg) c_dt = ldu(a.dt, p);
      if (!IRSMODE) c_TO = ldu(a.TO, p);
      if (c_m < 0) c_m = ldu(a.mat, p);
    }
    R n0 = 0, n4 = 0, n8 = 0;
    if (!TANG && live && k < k1) {
      n0 = ldu(a.W[0], p + sz);
      n4 = ldu(a.W[4], p + sz);
      n8 = ldu(a.W[8], p + sz);
    }
    R fz = 0;
    {
#pragma clang fp contract(off)
      if (live) {
        if (i < B.ni) fx[ty][tx + 1] = flux_x<TANG>(a, B, ring[sm], ring[sc], ring[sp], x, y, p, w0);
        if (j < B.nj) fy[ty + 1][tx] = flux_y<TANG>(a, B, ring[sm], ring[sc], ring[sp], x, y, p, w4);
        if (k < B.nk) fz = flux_z<TANG>(a, B, ring[sc], ring[sp], x, y, p, w8);
      }
      if (yhalo)
        fy[0][tx] = flux_y<TANG>(a, B, ring[sm], ring[sc], ring[sp], x, 0, p - sy, TANG ? (R)0 : ldu(a.W[4], p - sy));
      if (xhalo)
        fx[hr][0] = flux_x<TANG>(a, B, ring[sm], ring[sc], ring[sp], 0, hr + 1, ph, TANG ? (R)0 : ldu(a.W[0], ph));
    }
    w0 = n0;
    w4 = n4;
    w8 = n8;
    __syncthreads();
    if (live) {
      R Rv = 0;
      {
#pragma clang fp contract(off)
        if (i > 1) Rv = Rv + fx[ty][tx];
        if (i < B.ni) Rv = Rv - fx[ty][tx + 1];
        if (j > 1) Rv = Rv + fy[ty][tx];
        if (j < B.nj) Rv = Rv - fy[ty + 1][tx];
        if (k > 1) Rv = Rv + fz_prev;
        if (k < B.nk) Rv = Rv - fz;
      }
      Rv = apply_bc_faces(a, B, i, j, k, Rv);
      const R T = ring[sc][y][x];
      const int m = c_m;
      const R dt = a.use_dtg ? (R)(*a.dtg) : c_dt;
      int bad = 0;
      const R res = scaled_residual<PRIM>(a.tab, m, T, Rv, c_vol, c_qv, dt, &bad);
      if (IRSMODE) {
        stu(a.Rres, p, res);
      } else {
        const R TO = c_TO;
        const R Tn = new_state<PRIM>(a.tab, m, T, TO, res, a.rkc, a.jam, &bad);
        stu(a.Tout, p, Tn);
        if (Tn != Tn || Tn < (R)0) flag_state(a.status, a.irk, t.b, i, j, k);
        const double d = (double)Tn - (double)TO;
        rsq += d * d;
      }
      if (bad) flag_error(a.status, FUSS_ETABLE, t.b, i, j, k);
    }
    fz_prev = fz;
    const int s0 = sm;
    sm = sc;
    sc = sp;
    sp = s0;
  }
  if (!IRSMODE && a.last) block_sum(rsq, a.partials, wg);
}


// ------------------------------------------------------------------------
// Software-pipelined 2.5-D stage kernel for orthogonal 3-D blocks (the
// normal face weights only).  Per plane the table rows of the plane's faces
// and cell are requested first (their addresses come from T in LDS), then
// the loads of the next plane (T of plane k+2 and the per-cell streams of
// k+1), then the flux arithmetic waits for the table rows only: loads retire
// in issue order, so the next plane's streams stay in flight across the
// whole plane.  Every load in the loop is unconditional (out-of-block lanes
// and the last plane read valid cells they then ignore), so the compiler's
// wait counts can leave the later loads outstanding.  UM: every block of the
// level has one material (uniform row base, no material stream).
// Same operations and order as k_stage3d<false, ...>.
// ------------------------------------------------------------------------
template <typename R>
struct StageStream {
  R vol, qv, dt, TO;   // per-cell streams
  R w0, w4, w8;        // normal weights of the cell's +x, +y, +z faces
  R wh;                // halo lanes: weight of the tile's min-side face
  int m, mh;           // material of the cell / of the halo face's left cell
};

// fetch_plane with every load issued: halo cells beyond the ghost layer read
// the tile's corner cell instead (a ghost or interior cell of the block); no
// flux of an in-block face uses them
template <int TY, typename R>
__device__ __forceinline__ void fetch_plane_u(PlaneRegs<R>& r, const R* __restrict__ T, const BlockInfo& B, int i0,
                                              int j0, int kk) {
  constexpr int TT = TILE_X * TY;
  static_assert(PX * (TY + 2) <= 2 * TT, "two values per thread");
  const unsigned corner = (unsigned)(B.base + (i0 - 2) + (long)(j0 - 2) * B.sy + (long)(kk - 1) * B.sz);
  {
    const int q = threadIdx.x, ii = q % PX, jj = q / PX;
    const bool in = i0 - 1 + ii <= B.ni + 1 && j0 - 1 + jj <= B.nj + 1;
    r.v0 = ldu(T, in ? corner + ii + jj * (unsigned)B.sy : corner);
  }
  {
    const int q = threadIdx.x + TT, ii = q % PX, jj = q / PX;
    const bool in = q < PX * (TY + 2) && i0 - 1 + ii <= B.ni + 1 && j0 - 1 + jj <= B.nj + 1;
    r.v1 = ldu(T, in ? corner + ii + jj * (unsigned)B.sy : corner);
  }
}

// TY: rows of the column (4 or 8, one wave per row; the 8-row column
// evaluates 9 instead of 10 y faces and loads 10 instead of 12 T rows per 8
// rows of cells)
template <bool UM, bool PRIM, bool IRSMODE, bool RECT, bool DTC, typename R, int KCT, int TY>
__global__ __launch_bounds__(TILE_X * TY) void k_stage3p(StageArgs<R> a) {
  constexpr int TT = TILE_X * TY;
  if (a.halt && *a.halt) return;
  const int wg = xcd_remap(blockIdx.x, gridDim.x);
  const Tile t = a.ctiles[wg];
  const BlockInfo B = a.blk[t.b];
  static_assert(TILE_X == 64, "one wave per tile row");
  // the row of the lane's wave: uniform, so tests on it are scalar branches
  const int tx = threadIdx.x % TILE_X, ty = __builtin_amdgcn_readfirstlane(threadIdx.x / TILE_X);
  const int i = t.i0 + tx, j = t.j0 + ty;
  const int k0 = t.k0, k1 = min(B.nk, k0 + KCT - 1);
  const bool live = i <= B.ni && j <= B.nj;
  const int x = tx + 1, y = ty + 1;
  // Rectilinear levels evaluate every face of the tile, the block's faces
  // with weight 0 (per-axis factors f = 0 there): their fluxes are +-0 and
  // are added unconditionally (the sum starts at +0 and never becomes -0,
  // so adding +-0 changes nothing)
  const bool yhalo = ty == 0 && (RECT || t.j0 > 1) && i <= B.ni;
  const int hr = tx;
  // (rectilinear levels: also where the column starts the block, with weight 0)
  const bool xhalo = ty == 1 && tx < TY && (RECT || t.i0 > 1) && t.j0 + hr <= B.nj;
  const bool hx = i < B.ni, hy = j < B.nj;  // the cell's +x / +y face is interior
  // the halo face is interior (table errors are counted for these only)
  const bool yhalo_in = yhalo && t.j0 > 1, xhalo_in = xhalo && t.i0 > 1;
  const int hq = xhalo ? hr + 1 : 1;        // LDS row of the x-halo face (in range for every lane)
  __shared__ R ring[3][TY + 2][PX];
  __shared__ R fx[TY][TILE_X + 1];
  __shared__ R fy[TY + 1][TILE_X];
  using V2 = typename Vec2<R>::T;
  using V4 = typename Vec4<R>::T;
  const int Ti1 = a.tab.Ti1, Ti2 = a.tab.Ti2, nT = a.tab.nT;
  const int Ti2m1 = to_vgpr(Ti2 - 1);
  const int mu = B.umat >= 0 ? B.umat : 0;
  // a material's rows addressed by the row number itself (row_sat)
  const V2* kap0 = reinterpret_cast<const V2*>(a.tab.kap2) - Ti1;
  const V4* pr0 = reinterpret_cast<const V4*>(a.tab.pr4) - Ti1;
  const V2* kap_u = kap0 + (long)mu * nT;
  const V4* pr_u = pr0 + (long)mu * nT;
  auto kin = [&](R T) -> int { return !(T >= (R)Ti1 && T < (R)Ti2); };
  // the halo face's weight array: +y weights for the y-halo row, +x for the x-halo
  const R* __restrict__ Wh = yhalo ? a.W[4] : a.W[0];
  // RECT (every block rectilinear): volume and normal weights are products of
  // per-axis factors, the lane's (i, j) part formed once, the column's k
  // parts staged in LDS (a global load inside the plane loop would make the
  // in-order wait for it also wait for the next plane's prefetch)
  R cvol = 0, cw0 = 0, cw4 = 0, cwh = 0;
  // a_z, f_z (and 1/dl_z^2 for DTC) of planes k0-1 .. k0+KCT-1
  __shared__ double zfac[DTC ? 3 : 2][KCT + 1];
  double qxy = 0.0;  // DTC: 1/dl_x^2 + 1/dl_y^2 of the lane's column
  if (RECT) {
    const RectAxes rx_ = rect_axes(a.rect, B);
    if (threadIdx.x < (DTC ? 3 : 2) * (KCT + 1)) {
      const int q = threadIdx.x % (KCT + 1), kk = min(k0 - 1 + q, B.nk + 1), row = threadIdx.x / (KCT + 1);
      // (f_z = 0 on the block's last plane: its +z flux is +-0)
      zfac[row][q] = row == 0 ? rx_.az[kk] : (row == 1 ? (kk < B.nk ? rx_.fz[kk] : 0.0) : rx_.qz[kk]);
    }
    const int ci_ = live ? i : 1, cj_ = live ? j : 1;
    cvol = (R)(rx_.ax[ci_] * rx_.ay[cj_]);
    // (0 on the block's +x face: the face's flux is then +-0 and is added
    // unconditionally)
    cw0 = hx ? (R)(rx_.fx[ci_] * rx_.ay[cj_]) : (R)0;
    if (DTC) qxy = rx_.qx[ci_] + rx_.qy[cj_];
    cw4 = hy ? (R)(rx_.fy[cj_] * rx_.ax[ci_]) : (R)0;
    // the halo face: +y face of (i, j0-1) or +x face of (i0-1, j0+hr)
    cwh = yhalo_in ? (R)(rx_.fy[t.j0 - 1] * rx_.ax[i])
                   : (xhalo_in ? (R)(rx_.fx[t.i0 - 1] * rx_.ay[t.j0 + hr]) : (R)0);
  }
  // Steady primitive stages with IRS on levels of one material per block
  // (W): the table rows of the column's temperatures are staged in LDS, a
  // window of TWR rows around the first plane's rows, and every lookup reads
  // LDS instead of memory (no table load in the in-order counter, so the
  // waits for the rows no longer wait for the next plane's streams).  A
  // lookup of a face or cell outside the window counts as a miss; a
  // work-group with a miss computes its column again from memory (its
  // stores are overwritten; the same operations give the same bits).
  constexpr bool WIN = UM && PRIM && IRSMODE;
  constexpr int TWR = 128;
  __shared__ V2 wkap[WIN ? TWR : 1];
  __shared__ V4 wpr[WIN ? TWR : 1];
  __shared__ int wred[2][TY];
  double rsq = 0.0, dtmin = 1e300;
  auto column = [&](auto wtag) -> bool {
    constexpr bool W = decltype(wtag)::value;
    int sm = 0, sc = 1, sp = 2;
    PlaneRegs<R> pre;
    fetch_plane_u<TY>(pre, a.T, B, t.i0, t.j0, k0 - 1);
    store_plane<TY>(ring[sm], pre);
    fetch_plane_u<TY>(pre, a.T, B, t.i0, t.j0, k0);
    store_plane<TY>(ring[sc], pre);
    fetch_plane_u<TY>(pre, a.T, B, t.i0, t.j0, k0 + 1);
    const unsigned sz = (unsigned)B.sz;
    // addresses of plane k0: the cell (cell (1,1) for out-of-block lanes) and the
    // halo face's left cell (the cell itself for other lanes)
    const int ci = live ? i : 1, cj = live ? j : 1;
    unsigned p = (unsigned)(B.base + (ci - 1) + (long)(cj - 1) * B.sy + (long)(k0 - 1) * B.sz);
    // (where the x-halo face is a block face, the row's own first cell: the
    // material read for it is then valid)
    unsigned ph = yhalo ? (t.j0 > 1 ? p - (unsigned)B.sy : p)
                        : (xhalo ? (unsigned)(B.base + max(t.i0 - 2, 0) + (long)(t.j0 + hr - 1) * B.sy +
                                              (long)(k0 - 1) * B.sz)
                                 : p);
    // face-map positions of the lane's boundary faces (BC values in face
    // order; at most one face per axis since every block dimension is >= 2)
    const int fa = live ? B.face_active : 0;
    const int xo = (fa & 1) && i == 1 ? B.fmap_off[0] + (j - 1) : ((fa & 2) && i == B.ni ? B.fmap_off[1] + (j - 1) : -1);
    const int yo = (fa & 4) && j == 1 ? B.fmap_off[2] + (i - 1) : ((fa & 8) && j == B.nj ? B.fmap_off[3] + (i - 1) : -1);
    const int zlo = (fa & 16) ? B.fmap_off[4] + (i - 1) + B.ni * (j - 1) : -1;
    const int zhi = (fa & 32) ? B.fmap_off[5] + (i - 1) + B.ni * (j - 1) : -1;
    auto zo_of = [&](int kk) { return kk == 1 ? zlo : (kk == B.nk ? zhi : -1); };
    // which loads of boundary values a tile / row needs at all
    const int fa_t = B.face_active;
    const bool xtile = (fa_t & 3) && (t.i0 == 1 || t.i0 + TILE_X - 1 >= B.ni);
    const bool yrow = (fa_t & 12) && (j == 1 || j == B.nj);
    // byte offsets of the lane's BC values in bcf (x and y: base + step per
    // plane); lanes without an entry read the entry that stays 0, so a tile,
    // row or plane with boundary faces subtracts its values unconditionally
    constexpr unsigned SR = sizeof(R);
    const unsigned xb8 = (xo >= 0 ? (unsigned)xo : a.bcz) * SR, xs8 = xo >= 0 ? (unsigned)B.nj * SR : 0u;
    const unsigned yb8 = (yo >= 0 ? (unsigned)yo : a.bcz) * SR, ys8 = yo >= 0 ? (unsigned)B.ni * SR : 0u;
    const unsigned zl8 = (zlo >= 0 ? (unsigned)zlo : a.bcz) * SR, zh8 = (zhi >= 0 ? (unsigned)zhi : a.bcz) * SR;
    const R dtg = a.use_dtg ? (R)(*a.dtg) : (R)0;
    // a block of uniform qvol: every lane reads the value of cell (1,1,1) (one
    // cached line instead of a stream)
    const unsigned quni = (unsigned)B.base;
    auto stream = [&](StageStream<R>& s, unsigned pp, unsigned hh) {
      s.qv = ldu(a.qvol, B.uqv ? quni : pp);
      if (!DTC) s.dt = ldn(a.dt, pp);
      if (!IRSMODE) s.TO = ldn(a.TO, pp);
      if (!RECT) {
        s.vol = ldn(a.vol, pp);
        s.w0 = ldn(a.W[0], pp);
        s.w4 = ldn(a.W[4], pp);
        s.w8 = ldn(a.W[8], pp);
        s.wh = ldu(Wh, hh);
      }
      if (!UM) {
        s.m = ldu(a.mat, pp);
        s.mh = ldu(a.mat, hh);
      }
    };
    StageStream<R> sa_, sb_;
    sa_.m = sa_.mh = sb_.m = sb_.mh = mu;
    sa_.TO = sb_.TO = (R)0;
    if (DTC) sa_.dt = sb_.dt = (R)0;
    if (RECT) sa_.vol = sb_.vol = sa_.w0 = sb_.w0 = sa_.w4 = sb_.w4 = sa_.w8 = sb_.w8 = sa_.wh = sb_.wh = (R)0;
    stream(sa_, p, ph);
    __syncthreads();
    // the window: TWR rows centred on the rows of plane k0's cells (every
    // work-group of a level of one material per block reads its block's rows)
    int wlo = Ti1;
    // (lanes with a miss, or-ed over the planes: a uniform mask)
    unsigned long long miss = 0;
    if constexpr (W) {
      const int rq = row_sat(ring[sc][y][x], Ti1, Ti2m1);
      // (sentinels for lanes outside the block: their difference cannot overflow)
      int rlo = live ? rq : 0x3fffffff, rhi = live ? rq : -0x3fffffff;
      for (int o = 32; o > 0; o >>= 1) {
        rlo = min(rlo, __shfl_xor(rlo, o, 64));
        rhi = max(rhi, __shfl_xor(rhi, o, 64));
      }
      if (tx == 0) {
        wred[0][ty] = rlo;
        wred[1][ty] = rhi;
      }
      __syncthreads();
      int lo_ = wred[0][0], hi_ = wred[1][0];
      for (int q = 1; q < TY; q++) {
        lo_ = min(lo_, wred[0][q]);
        hi_ = max(hi_, wred[1][q]);
      }
      // (a plane spanning more than half the window: the memory path)
      if (hi_ - lo_ >= TWR / 2) return false;
      wlo = __builtin_amdgcn_readfirstlane(max(Ti1, min((lo_ + hi_) / 2 - TWR / 2, Ti2 - TWR)));
      for (int q = threadIdx.x; q < TWR; q += TT) {
        wkap[q] = kap_u[wlo + q];
        wpr[q] = pr_u[wlo + q];
      }
      __syncthreads();
    }
    const int wtop = __builtin_amdgcn_readfirstlane(TWR - 1);
    // a row clamped to the table, then its place in the window clamped to
    // the window; ms: the two differ
    auto wrow = [&](R Tv, int& ms) {
      const int r = row_sat(Tv, Ti1, Ti2m1) - wlo;
      const int rw = med3_i32_0s(r, wtop);
      ms = r != rw;
      return rw;
    };
    R fz_prev = 0;
    if (live && k0 > 1) {
#pragma clang fp contract(off)
      fz_prev = flux_z<false>(a, B, ring[sm], ring[sc], x, y, p - sz,
                              RECT ? (R)(zfac[1][0] * (double)cvol) : ldu(a.W[8], p - sz));
    }
    // The result of a plane is stored during the next plane, after that
    // plane's loads are issued: stores count in the same in-order counter as
    // loads, so a store issued just before the table loads would make the
    // wait for the tables wait for the store's completion too.
    R* const out = IRSMODE ? a.Rres : a.Tout;
    R vpend = 0, dtpend = 0;
    unsigned ppend = p;
    bool spend = false;
    // one plane: cur holds its streams, nxt receives those of the next plane
    auto plane = [&](StageStream<R>& cur, StageStream<R>& nxt, int k) {
      store_plane<TY>(ring[sp], pre);
      __syncthreads();
      // phase A: table rows of the plane's faces and of the cell (loads always
      // issued, at a clamped row; errors counted for existing faces only)
      int bad = 0, badc = 0;
      const R Tc = ring[sc][y][x];
      const R Txr = ring[sc][y][x + 1], Tyr = ring[sc][y + 1][x], Tzr = ring[sp][y][x];
      // face averages, then the fractions of their table rows (T - row, the
      // row taken as a floating value: no conversion back from the integer)
      R Tx, Ty, Tz, Th = 0, Tcl = Tc, fc = 0;
      V2 kx, ky, kz, kh{}, kc{};
      V4 pc{};
      // BC values of the lane's boundary faces, issued with the table rows;
      // only tiles on an i face, rows on a j face and the end planes load them
      // (each read only under the test that loads it)
      R bx, by, bz;
      // (a lane's z face has an entry only on an end plane of a block with z entries)
      const bool zpl = ((fa_t & 16) && k == 1) || ((fa_t & 32) && k == B.nk);
      if (xtile) bx = ldb<R>(a.bcf, xb8 + xs8 * (unsigned)(k - 1));
      if (yrow) by = ldb<R>(a.bcf, yb8 + ys8 * (unsigned)(k - 1));
      if (zpl) bz = ldb<R>(a.bcf, k == 1 ? zl8 : zh8);
      {
#pragma clang fp contract(off)
        // rows addressed by 32-bit offsets from the material's rows; the
        // fraction T - row is x - floor(x) (the same for T in the table; outside
        // it an error is flagged)
        const V2* kr = UM ? kap_u : kap0 + (long)cur.m * nT;
        int b0 = 0, b1 = 0, b2 = 0, b3 = 0;
        int mx = 0, my = 0, mz = 0, mc = 0, mh = 0;
        Tx = (R)0.5 * (Tc + Txr);
        if constexpr (W) kx = wkap[wrow(Tx, mx)];
        else kx = ldu(kr, (unsigned)row_sat(Tx, Ti1, Ti2m1));
        if (!PRIM) b0 = kin(Tx);
        Tx = frac_of(Tx);
        Ty = (R)0.5 * (Tc + Tyr);
        if constexpr (W) ky = wkap[wrow(Ty, my)];
        else ky = ldu(kr, (unsigned)row_sat(Ty, Ti1, Ti2m1));
        if (!PRIM) b1 = kin(Ty);
        Ty = frac_of(Ty);
        Tz = (R)0.5 * (Tc + Tzr);
        if constexpr (W) kz = wkap[wrow(Tz, mz)];
        else kz = ldu(kr, (unsigned)row_sat(Tz, Ti1, Ti2m1));
        if (!PRIM) b2 = kin(Tz);
        Tz = frac_of(Tz);
        if (PRIM) {
          if constexpr (W) {
            const int qc = wrow(Tcl, mc);
            pc = wpr[qc];
            if (DTC) kc = wkap[qc];
          } else {
            const int rc = row_sat(Tcl, Ti1, Ti2m1);
            pc = ldu(UM ? pr_u : pr0 + (long)cur.m * nT, (unsigned)rc);
            if (DTC) kc = ldu(kr, (unsigned)rc);  // kappa at the cell for its dt
          }
          badc = kin(Tcl);
          fc = frac_of(Tcl);
        }
        if (ty < 2) {  // waves 0 and 1 hold the halo lanes (a wave-uniform test)
          const V2* krh = UM ? kap_u : kap0 + (long)cur.mh * nT;
          Th = yhalo ? (R)0.5 * (ring[sc][0][x] + ring[sc][1][x]) : (R)0.5 * (ring[sc][hq][0] + ring[sc][hq][1]);
          if constexpr (W) kh = wkap[wrow(Th, mh)];
          else kh = ldu(krh, (unsigned)row_sat(Th, Ti1, Ti2m1));
          if (!PRIM) b3 = kin(Th);
          Th = frac_of(Th);
        }
        // primitive variables: every cell's own row is checked (badc), and a face
        // average lies between its two cells' temperatures, so a face outside
        // the table implies a cell outside it -- the face tests are redundant
        bad = PRIM ? 0 : (live && ((hx && b0) || (hy && b1) || (k < B.nk && b2))) || ((yhalo_in || xhalo_in) && b3);
        badc = live && badc;
        // a row of a face whose flux is used (or of the cell) outside the window
        if constexpr (W)
          miss |= __builtin_amdgcn_ballot_w64((live && ((hx && mx) || (hy && my) || (k < B.nk && mz) || mc)) ||
                                              ((yhalo_in || xhalo_in) && mh));
      }
      // phase B: the next plane's loads (the last plane re-reads its own cells)
      {
        const unsigned step = k < k1 ? sz : 0u;
        fetch_plane_u<TY>(pre, a.T, B, t.i0, t.j0, k < k1 ? k + 2 : k + 1);
        stream(nxt, p + step, ph + step);
      }
      if (spend) {  // the previous plane's result
        stn(out, ppend, vpend);
        if (DTC) stn(a.dtw, ppend, dtpend);
      }
      if (RECT) {
        // the plane's geometry from the per-axis factors (same products as the
        // setup check against the stored metrics)
        const double azk = zfac[0][k - k0 + 1], fzk = zfac[1][k - k0 + 1];
        cur.vol = (R)((double)cvol * azk);
        cur.w0 = (R)((double)cw0 * azk);
        cur.w4 = (R)((double)cw4 * azk);
        cur.w8 = (R)(fzk * (double)cvol);
        cur.wh = (R)((double)cwh * azk);
      }
      // phase C: fluxes
      R fz = 0;
      {
#pragma clang fp contract(off)
        if (live) {
          if (RECT || hx) fx[ty][tx + 1] = (kx.x + Tx * kx.y) * ((Txr - Tc) * cur.w0);
          if (RECT || hy) fy[ty + 1][tx] = (ky.x + Ty * ky.y) * ((Tyr - Tc) * cur.w4);
          if (RECT || k < B.nk) fz = (kz.x + Tz * kz.y) * ((Tzr - Tc) * cur.w8);
        }
        if (yhalo) fy[0][tx] = (kh.x + Th * kh.y) * ((ring[sc][1][x] - ring[sc][0][x]) * cur.wh);
        if (xhalo) fx[hr][0] = (kh.x + Th * kh.y) * ((ring[sc][hq][1] - ring[sc][hq][0]) * cur.wh);
      }
      if (bad) flag_error(a.status, FUSS_ETABLE, -1, 0, 0, 0);
      __syncthreads();
      if (live) {
        R Rv = 0;
        {
#pragma clang fp contract(off)
          // (rectilinear levels: the block's faces hold +-0; fz_prev is +0 on
          // the block's first plane)
          if (RECT || i > 1) Rv = Rv + fx[ty][tx];
          if (RECT || hx) Rv = Rv - fx[ty][tx + 1];
          if (RECT || j > 1) Rv = Rv + fy[ty][tx];
          if (RECT || hy) Rv = Rv - fy[ty + 1][tx];
          if (RECT || k > 1) Rv = Rv + fz_prev;
          if (RECT || k < B.nk) Rv = Rv - fz;
        }
        {
#pragma clang fp contract(off)
          // BC_Fluxes contributions in face order (as apply_bc_faces); +0
          // where the cell has none
          // (scalar branches: an opaque statement keeps them from being turned
          // into selects)
          if (xtile) {
            keep_branch();
            Rv = Rv - bx;
          }
          if (yrow) {
            keep_branch();
            Rv = Rv - by;
          }
          if (zpl) {
            keep_branch();
            Rv = Rv - bz;
          }
        }
        R dt = a.use_dtg ? dtg : cur.dt;
        if (DTC) {
          // Compute_dt (Mod_dt.f90:26-84) at T^n, the first stage's input: the
          // kappa, rho, cp interpolations of k_dt at the cell's row (rho and cp
          // are the residual's below), sum 1/dl^2 from the per-axis factors
#pragma clang fp contract(off)
          const R fr = fc;
          const R kap = kc.x + fr * kc.y, rho = pc.x + fr * pc.y, cp = pc.z + fr * pc.w;
          const double alpha = (double)kap / (double)rho / (double)cp;
          double dd = 0.5 * a.vnn / (alpha * (qxy + zfac[2][k - k0 + 1]));
          dd = dd < 1e8 ? dd : 1e8;
          dt = (R)dd;
          dtpend = dt;
          dtmin = fmin(dtmin, dd);
        }
        R res;
        {
#pragma clang fp contract(off)
          if (PRIM) {
            // Compute_Prim_Residual, Lib_Newstate.f90:241-256 (as scaled_residual)
            const R rho = pc.x + fc * pc.y;
            const R cp = pc.z + fc * pc.w;
            res = (-Rv / cur.vol + cur.qv) * dt / (Tc * cp * pc.y + rho * cp + Tc * rho * pc.w);
          } else {
            res = (-Rv / cur.vol + cur.qv) * dt;
          }
        }
        int bad2 = badc;
        if (IRSMODE) {
          vpend = res;
        } else {
          const R Tn = new_state<PRIM>(a.tab, cur.m, Tc, cur.TO, res, a.rkc, a.jam, &bad2);
          vpend = Tn;
          if (Tn != Tn || Tn < (R)0) flag_state(a.status, a.irk, t.b, i, j, k);
          const double d = (double)Tn - (double)cur.TO;
          rsq += d * d;
        }
        ppend = p;
        if (bad2) flag_error(a.status, FUSS_ETABLE, t.b, i, j, k);
      }
      spend = live;
      fz_prev = fz;
      const int s0 = sm;
      sm = sc;
      sc = sp;
      sp = s0;
      p += sz;
      ph += sz;
    };
    // two planes per trip with the stream buffers swapping roles (no copy of
    // in-flight registers)
    for (int k = k0; k <= k1; k += 2) {
      plane(sa_, sb_, k);
      if (k + 1 <= k1) plane(sb_, sa_, k + 1);
    }
    if (spend) {
      stn(out, ppend, vpend);
      if (DTC) stn(a.dtw, ppend, dtpend);
    }
    // the column is done unless a lane of the work-group missed the window
    if constexpr (W) return !__syncthreads_or(miss != 0);
    return true;
  };
  bool done = false;
  if constexpr (WIN) {
    if (Ti2 - Ti1 >= TWR) done = column(std::true_type{});
  }
  if (!done) {
    rsq = 0.0;
    dtmin = 1e300;
    column(std::false_type{});
  }
  if (!IRSMODE && a.last) block_sum<TT>(rsq, a.partials, wg);
  if (DTC) block_min<TT>(dtmin, a.pmin, wg);
}

// ------------------------------------------------------------------------
// Software-pipelined stage kernel for 3-D blocks with tangential weights
// (curvilinear grids): k_stage3d's arithmetic on k_stage3p's schedule.  Per
// plane the table rows of the plane's faces and cell are requested first,
// then T of plane k+2 and the per-cell streams of plane k+1 (1/V, qvol, dt,
// T^n and the nine face weights, plus the three weights of the lane's halo
// face), so the streams stay in flight across a whole plane; every load in
// the loop is unconditional.  The tangential gradients read the planes k-1,
// k, k+1 of the LDS ring before the plane's second barrier, so the slot of
// k-1 is rewritten (with plane k+2) only after every wave is done with it.
// ------------------------------------------------------------------------
template <typename R>
struct StageStreamT {
  R vol, qv, dt, TO;
  int m, mh;  // material of the cell / of the halo face's left cell
};

template <bool UM, bool PRIM, bool IRSMODE, typename R, int KCT>
// (three waves per SIMD: the FP64 instantiations then fit 168 VGPRs with one
// value spilled to scratch per two planes; the compiler's own choice, 175
// VGPRs and two waves, runs the lattice's stage 1.3 % slower)
__global__ __launch_bounds__(TILE_T) __attribute__((amdgpu_waves_per_eu(3))) void k_stage3t(StageArgs<R> a) {
  constexpr int TY = TILE_Y, TT = TILE_T;
  static_assert(TILE_X == 64, "one wave per tile row");
  if (a.halt && *a.halt) return;
  const int wg = xcd_remap(blockIdx.x, gridDim.x);
  const Tile t = a.ctiles[wg];
  const BlockInfo B = a.blk[t.b];
  const int tx = threadIdx.x % TILE_X, ty = __builtin_amdgcn_readfirstlane(threadIdx.x / TILE_X);
  const int i = t.i0 + tx, j = t.j0 + ty;
  const int k0 = t.k0, k1 = min(B.nk, k0 + KCT - 1);
  const bool live = i <= B.ni && j <= B.nj;
  const int x = tx + 1, y = ty + 1;
  const bool yhalo = ty == 0 && t.j0 > 1 && i <= B.ni;
  const int hr = tx;
  const bool xhalo = ty == 1 && tx < TY && t.i0 > 1 && t.j0 + hr <= B.nj;
  const bool hx = i < B.ni, hy = j < B.nj;
  const int hq = xhalo ? hr + 1 : 1;  // LDS row of the x-halo face (in range for every lane)
  __shared__ R ring[3][TY + 2][PX];
  __shared__ R fx[TY][TILE_X + 1];
  __shared__ R fy[TY + 1][TILE_X];
  using V2 = typename Vec2<R>::T;
  using V4 = typename Vec4<R>::T;
  const int Ti1 = a.tab.Ti1, Ti2 = a.tab.Ti2, nT = a.tab.nT;
  const int mu = B.umat >= 0 ? B.umat : 0;
  const V2* kap_u = reinterpret_cast<const V2*>(a.tab.kap2) + (long)mu * nT;
  const V4* pr_u = reinterpret_cast<const V4*>(a.tab.pr4) + (long)mu * nT;
  const V2* kap0 = reinterpret_cast<const V2*>(a.tab.kap2);
  const V4* pr0 = reinterpret_cast<const V4*>(a.tab.pr4);
  // the halo face's weights: +y face (W3, W4, W5) for the y-halo row, +x
  // (W0, W1, W2) for the x-halo lanes
  const int wbase = yhalo ? 3 : 0;
  int sm = 0, sc = 1, sp = 2;
  PlaneRegs<R> pre;
  fetch_plane_u<TY>(pre, a.T, B, t.i0, t.j0, k0 - 1);
  store_plane<TY>(ring[sm], pre);
  fetch_plane_u<TY>(pre, a.T, B, t.i0, t.j0, k0);
  store_plane<TY>(ring[sc], pre);
  fetch_plane_u<TY>(pre, a.T, B, t.i0, t.j0, k0 + 1);
  const unsigned sz = (unsigned)B.sz;
  const int ci = live ? i : 1, cj = live ? j : 1;
  unsigned p = (unsigned)(B.base + (ci - 1) + (long)(cj - 1) * B.sy + (long)(k0 - 1) * B.sz);
  unsigned ph = yhalo ? p - (unsigned)B.sy
                      : (xhalo ? (unsigned)(B.base + (t.i0 - 2) + (long)(t.j0 + hr - 1) * B.sy + (long)(k0 - 1) * B.sz)
                               : p);
  const int fa = live ? B.face_active : 0;
  const int xo = (fa & 1) && i == 1 ? B.fmap_off[0] + (j - 1) : ((fa & 2) && i == B.ni ? B.fmap_off[1] + (j - 1) : -1);
  const int yo = (fa & 4) && j == 1 ? B.fmap_off[2] + (i - 1) : ((fa & 8) && j == B.nj ? B.fmap_off[3] + (i - 1) : -1);
  const int zlo = (fa & 16) ? B.fmap_off[4] + (i - 1) + B.ni * (j - 1) : -1;
  const int zhi = (fa & 32) ? B.fmap_off[5] + (i - 1) + B.ni * (j - 1) : -1;
  auto zo_of = [&](int kk) { return kk == 1 ? zlo : (kk == B.nk ? zhi : -1); };
  const int fa_t = B.face_active;
  const bool xtile = (fa_t & 3) && (t.i0 == 1 || t.i0 + TILE_X - 1 >= B.ni);
  const bool yrow = (fa_t & 12) && (j == 1 || j == B.nj);
  const R dtg = a.use_dtg ? (R)(*a.dtg) : (R)0;
  const unsigned quni = (unsigned)B.base;
  // with a global dt every lane reads the dt of cell (1,1,1): no stream
  const bool dtu = a.use_dtg != 0;
  // the face weights (nine per cell, three for the halo face) are held for
  // one plane only: a plane's are requested after the previous plane's
  // fluxes, so they arrive while that plane finishes (registers for a second
  // set would halve the waves per SIMD)
  R w[9], wh[3];
  auto weights = [&](unsigned pp, unsigned hh) {
#pragma unroll
    for (int q = 0; q < 9; q++) w[q] = ldn(a.W[q], pp);
#pragma unroll
    for (int q = 0; q < 3; q++) wh[q] = ldu(a.W[wbase + q], hh);
  };
  auto stream = [&](StageStreamT<R>& s, unsigned pp, unsigned hh) {
    s.qv = ldu(a.qvol, B.uqv ? quni : pp);
    s.dt = ldn(a.dt, dtu ? quni : pp);
    if (!IRSMODE) s.TO = ldn(a.TO, pp);
    s.vol = ldn(a.vol, pp);
    if (!UM) {
      s.m = ldu(a.mat, pp);
      s.mh = ldu(a.mat, hh);
    }
  };
  StageStreamT<R> sa_, sb_;
  sa_.m = sa_.mh = sb_.m = sb_.mh = mu;
  sa_.TO = sb_.TO = (R)0;
  stream(sa_, p, ph);
  weights(p, ph);
  __syncthreads();
  R fz_prev = 0;
  if (live && k0 > 1) {
#pragma clang fp contract(off)
    fz_prev = flux_z<true>(a, B, ring[sm], ring[sc], x, y, p - sz, (R)0);
  }
  double rsq = 0.0;
  R* const out = IRSMODE ? a.Rres : a.Tout;
  R vpend = 0;
  unsigned ppend = p;
  bool spend = false;
  auto plane = [&](StageStreamT<R>& cur, StageStreamT<R>& nxt, int k) {
    store_plane<TY>(ring[sp], pre);
    __syncthreads();
    // phase A: table rows of the plane's faces and of the cell
    int bad = 0, badc = 0;
    const R Tc = ring[sc][y][x];
    const R Txr = ring[sc][y][x + 1], Tyr = ring[sc][y + 1][x], Tzr = ring[sp][y][x];
    R Tx, Ty, Tz, Th = 0, Tcl = Tc, fc = 0;
    V2 kx, ky, kz, kh{};
    V4 pc{};
    R bx = 0, by = 0, bz = 0;
    if (xtile) bx = ldu(a.bcf, (unsigned)max(xo >= 0 ? xo + B.nj * (k - 1) : 0, 0));
    if (yrow) by = ldu(a.bcf, (unsigned)max(yo >= 0 ? yo + B.ni * (k - 1) : 0, 0));
    if ((fa_t & 48) && (k == 1 || k == B.nk)) bz = ldu(a.bcf, (unsigned)max(zo_of(k), 0));
    {
#pragma clang fp contract(off)
      const V2* kr = UM ? kap_u : kap0 + (long)cur.m * nT;
      int b0 = 0, b1 = 0, b2 = 0, b3 = 0;
      Tx = (R)0.5 * (Tc + Txr);
      const R rxf = tab_rowf(Tx, Ti1, Ti2, &b0);
      kx = ldu(kr, (unsigned)((int)rxf - Ti1));
      Tx = Tx - rxf;
      Ty = (R)0.5 * (Tc + Tyr);
      const R ryf = tab_rowf(Ty, Ti1, Ti2, &b1);
      ky = ldu(kr, (unsigned)((int)ryf - Ti1));
      Ty = Ty - ryf;
      Tz = (R)0.5 * (Tc + Tzr);
      const R rzf = tab_rowf(Tz, Ti1, Ti2, &b2);
      kz = ldu(kr, (unsigned)((int)rzf - Ti1));
      Tz = Tz - rzf;
      if (PRIM) {
        const R rcf = tab_rowf(Tcl, Ti1, Ti2, &badc);
        pc = ldu(UM ? pr_u : pr0 + (long)cur.m * nT, (unsigned)((int)rcf - Ti1));
        fc = Tcl - rcf;
      }
      if (ty < 2) {  // waves 0 and 1 hold the halo lanes
        const V2* krh = UM ? kap_u : kap0 + (long)cur.mh * nT;
        Th = yhalo ? (R)0.5 * (ring[sc][0][x] + ring[sc][1][x]) : (R)0.5 * (ring[sc][hq][0] + ring[sc][hq][1]);
        const R rhf = tab_rowf(Th, Ti1, Ti2, &b3);
        kh = ldu(krh, (unsigned)((int)rhf - Ti1));
        Th = Th - rhf;
      }
      // every face inside the block is checked, like k_stage3d's flux evaluations
      bad = (live && ((hx && b0) || (hy && b1) || (k < B.nk && b2))) || ((yhalo || xhalo) && b3);
      badc = live && badc;
    }
    // phase B: the next plane's loads (the last plane re-reads its own cells)
    {
      const unsigned step = k < k1 ? sz : 0u;
      fetch_plane_u<TY>(pre, a.T, B, t.i0, t.j0, k < k1 ? k + 2 : k + 1);
      stream(nxt, p + step, ph + step);
    }
    if (spend) stn(out, ppend, vpend);
    // phase C: fluxes (Diffusive_Flux regrouped per direction, as flux_x/y/z)
    R (*m_)[PX] = ring[sm];
    R (*c_)[PX] = ring[sc];
    R (*p_)[PX] = ring[sp];
    R fz = 0;
    {
#pragma clang fp contract(off)
      if (live) {
        if (hx) {
          const R G0 = Txr - Tc;
          const R G1 = (c_[y + 1][x] - c_[y - 1][x] + c_[y + 1][x + 1] - c_[y - 1][x + 1]) * (R)0.25;
          const R G2 = (p_[y][x] - m_[y][x] + p_[y][x + 1] - m_[y][x + 1]) * (R)0.25;
          fx[ty][tx + 1] = (kx.x + Tx * kx.y) * ((G0 * w[0] + G1 * w[1]) + G2 * w[2]);
        }
        if (hy) {
          const R G1 = Tyr - Tc;
          const R G0 = (c_[y][x + 1] - c_[y][x - 1] + c_[y + 1][x + 1] - c_[y + 1][x - 1]) * (R)0.25;
          const R G2 = (p_[y][x] - m_[y][x] + p_[y + 1][x] - m_[y + 1][x]) * (R)0.25;
          fy[ty + 1][tx] = (ky.x + Ty * ky.y) * ((G0 * w[3] + G1 * w[4]) + G2 * w[5]);
        }
        if (k < B.nk) {
          const R G2 = Tzr - Tc;
          const R G0 = (c_[y][x + 1] - c_[y][x - 1] + p_[y][x + 1] - p_[y][x - 1]) * (R)0.25;
          const R G1 = (c_[y + 1][x] - c_[y - 1][x] + p_[y + 1][x] - p_[y - 1][x]) * (R)0.25;
          fz = (kz.x + Tz * kz.y) * ((G0 * w[6] + G1 * w[7]) + G2 * w[8]);
        }
      }
      if (yhalo) {  // +y face of (i, j0-1): LDS row 0
        const R G1 = c_[1][x] - c_[0][x];
        const R G0 = (c_[0][x + 1] - c_[0][x - 1] + c_[1][x + 1] - c_[1][x - 1]) * (R)0.25;
        const R G2 = (p_[0][x] - m_[0][x] + p_[1][x] - m_[1][x]) * (R)0.25;
        fy[0][tx] = (kh.x + Th * kh.y) * ((G0 * wh[0] + G1 * wh[1]) + G2 * wh[2]);
      }
      if (xhalo) {  // +x face of (i0-1, j0+hr): LDS column 0 of row hq
        const R G0 = c_[hq][1] - c_[hq][0];
        const R G1 = (c_[hq + 1][0] - c_[hq - 1][0] + c_[hq + 1][1] - c_[hq - 1][1]) * (R)0.25;
        const R G2 = (p_[hq][0] - m_[hq][0] + p_[hq][1] - m_[hq][1]) * (R)0.25;
        fx[hr][0] = (kh.x + Th * kh.y) * ((G0 * wh[0] + G1 * wh[1]) + G2 * wh[2]);
      }
    }
    if (bad) flag_error(a.status, FUSS_ETABLE, -1, 0, 0, 0);
    {
      const unsigned step = k < k1 ? sz : 0u;
      weights(p + step, ph + step);
    }
    __syncthreads();
    if (live) {
      R Rv = 0;
      {
#pragma clang fp contract(off)
        if (i > 1) Rv = Rv + fx[ty][tx];
        if (hx) Rv = Rv - fx[ty][tx + 1];
        if (j > 1) Rv = Rv + fy[ty][tx];
        if (hy) Rv = Rv - fy[ty + 1][tx];
        if (k > 1) Rv = Rv + fz_prev;
        if (k < B.nk) Rv = Rv - fz;
      }
      {
#pragma clang fp contract(off)
        if (xo >= 0) Rv = Rv - bx;
        if (yo >= 0) Rv = Rv - by;
        if (zo_of(k) >= 0) Rv = Rv - bz;
      }
      const R dt = a.use_dtg ? dtg : cur.dt;
      R res;
      {
#pragma clang fp contract(off)
        if (PRIM) {
          const R rho = pc.x + fc * pc.y;
          const R cp = pc.z + fc * pc.w;
          res = (-Rv / cur.vol + cur.qv) * dt / (Tc * cp * pc.y + rho * cp + Tc * rho * pc.w);
        } else {
          res = (-Rv / cur.vol + cur.qv) * dt;
        }
      }
      int bad2 = badc;
      if (IRSMODE) {
        vpend = res;
      } else {
        const R Tn = new_state<PRIM>(a.tab, cur.m, Tc, cur.TO, res, a.rkc, a.jam, &bad2);
        vpend = Tn;
        if (Tn != Tn || Tn < (R)0) flag_state(a.status, a.irk, t.b, i, j, k);
        const double d = (double)Tn - (double)cur.TO;
        rsq += d * d;
      }
      ppend = p;
      if (bad2) flag_error(a.status, FUSS_ETABLE, t.b, i, j, k);
    }
    spend = live;
    fz_prev = fz;
    const int s0 = sm;
    sm = sc;
    sc = sp;
    sp = s0;
    p += sz;
    ph += sz;
  };
  for (int k = k0; k <= k1; k += 2) {
    plane(sa_, sb_, k);
    if (k + 1 <= k1) plane(sb_, sa_, k + 1);
  }
  if (spend) stn(out, ppend, vpend);
  if (!IRSMODE && a.last) block_sum<TT>(rsq, a.partials, wg);
}

// ------------------------------------------------------------------------
// Barrier-free stage kernel for orthogonal 3-D blocks: each wave owns one
// 64-cell row (i) of the tile and marches its KC planes in k on its own.
// i-neighbours come from the neighbouring lanes (shuffles; lane 0 and 63
// read the tile edges), j-neighbours from memory (the rows of the other
// waves, cache hits), the +z flux is carried to the next plane in a
// register.  The flux through a j face is evaluated by both rows that share
// it, from the same operands, so the residual is bit-identical to k_stage3d's.
// Everything a plane needs is fetched one plane ahead.
// ------------------------------------------------------------------------
template <typename R>
struct RowPlane {
  R T, Tjm, Tjp, Tih;   // own T, j-1 and j+1 neighbours, tile-edge halo (lanes 0, 63)
  R w0, w4, w8, w4m;    // normal weights of the cell's +x, +y, +z faces and of the (j-1) cell's +y face
  R w0h;                // lane 0: +x weight of cell i0-1
  R vol, qv, dt, TO;
  int m, mjm, mih;      // material of the cell, of (j-1), of (i0-1)
};

template <bool PRIM, bool IRSMODE, typename R>
__device__ __forceinline__ void row_fetch(RowPlane<R>& r, const StageArgs<R>& a, const BlockInfo& B, unsigned p,
                                          int lane, bool live, bool lo_halo, bool hi_halo, bool has_jm) {
  const unsigned sy = (unsigned)B.sy;
  if (live) {
    r.T = ldu(a.T, p);
    r.Tjm = ldu(a.T, p - sy);
    r.Tjp = ldu(a.T, p + sy);
    r.w0 = ldu(a.W[0], p);
    r.w4 = ldu(a.W[4], p);
    r.w8 = ldu(a.W[8], p);
    r.w4m = has_jm ? ldu(a.W[4], p - sy) : (R)0;
    r.vol = ldu(a.vol, p);
    r.qv = ldu(a.qvol, p);
    if (!a.use_dtg) r.dt = ldu(a.dt, p);
    if (!IRSMODE) r.TO = ldu(a.TO, p);
    if (B.umat < 0) {
      r.m = ldu(a.mat, p);
      r.mjm = has_jm ? ldu(a.mat, p - sy) : 0;
    }
  }
  if (lo_halo) {
    r.Tih = ldu(a.T, p - 1);
    r.w0h = ldu(a.W[0], p - 1);
    if (B.umat < 0) r.mih = ldu(a.mat, p - 1);
  } else if (hi_halo) {
    r.Tih = ldu(a.T, p + 1);
  }
  (void)lane;
}

template <typename R>
__device__ __forceinline__ R face_kw(const StageArgs<R>& a, int m, R Tl, R Tr, R w) {
  int bad = 0;
  const R kap = tab_interp2(a.tab.kap2 + 2 * (long)m * a.tab.nT, (R)0.5 * (Tl + Tr), a.tab.Ti1, a.tab.Ti2, &bad);
  if (bad) flag_error(a.status, FUSS_ETABLE, -1, 0, 0, 0);
  return kap * ((Tr - Tl) * w);
}

template <bool PRIM, bool IRSMODE, typename R, int KCT>
__global__ __launch_bounds__(TILE_T) void k_stage3w(StageArgs<R> a) {
  if (a.halt && *a.halt) return;
  const int wg = xcd_remap(blockIdx.x, gridDim.x);
  const Tile t = a.ctiles[wg];
  const BlockInfo B = a.blk[t.b];
  const int lane = threadIdx.x % TILE_X, row = threadIdx.x / TILE_X;
  const int i = t.i0 + lane, j = t.j0 + row;
  const int k0 = t.k0, k1 = min(B.nk, k0 + KCT - 1);
  const bool rowlive = j <= B.nj;
  const bool live = rowlive && i <= B.ni;
  const bool has_jm = j > 1;
  // lane 0 computes the tile's min-side x face; lane 63 reads the cell past the tile
  const bool lo_halo = rowlive && lane == 0 && t.i0 > 1;
  const bool hi_halo = rowlive && lane == TILE_X - 1 && i < B.ni;
  const unsigned sz = (unsigned)B.sz;
  unsigned p = (unsigned)(B.base + (i - 1) + (long)(j - 1) * B.sy + (long)(k0 - 1) * B.sz);
  RowPlane<R> cur, nxt;
  cur.m = cur.mjm = cur.mih = B.umat;
  nxt.m = nxt.mjm = nxt.mih = B.umat;
  cur.dt = nxt.dt = (R)0;
  cur.TO = nxt.TO = (R)0;
  cur.Tih = nxt.Tih = (R)0;
  cur.w0h = nxt.w0h = (R)0;
  row_fetch<PRIM, IRSMODE>(cur, a, B, p, lane, live, lo_halo, hi_halo, has_jm);
  R Tp = live && k0 < B.nk ? ldu(a.T, p + sz) : (R)0;
  R fz_prev = 0;
  if (live && k0 > 1) {
#pragma clang fp contract(off)
    const R Tm = ldu(a.T, p - sz);
    const int mm = B.umat >= 0 ? B.umat : ldu(a.mat, p - sz);
    fz_prev = face_kw(a, mm, Tm, cur.T, ldu(a.W[8], p - sz));
  }
  double rsq = 0.0;
  for (int k = k0; k <= k1; k++, p += sz) {
    // next plane's data (and T two planes up for the +z face of plane k+1)
    R Tp2 = 0;
    if (k < k1) {
      row_fetch<PRIM, IRSMODE>(nxt, a, B, p + sz, lane, live, lo_halo, hi_halo, has_jm);
      if (live && k + 1 < B.nk) Tp2 = ldu(a.T, p + 2 * sz);
    }
    // i neighbours from the neighbouring lanes
    R Tl = __shfl_up(cur.T, 1, 64), Tr = __shfl_down(cur.T, 1, 64);
    if (lane == 0) Tl = cur.Tih;
    if (lane == TILE_X - 1) Tr = cur.Tih;
    R fxp = 0, fyp = 0, fym = 0, fz = 0, fxm_own = 0;
    {
#pragma clang fp contract(off)
      if (live) {
        const int m = cur.m;
        if (i < B.ni) fxp = face_kw(a, m, cur.T, Tr, cur.w0);
        if (j < B.nj) fyp = face_kw(a, m, cur.T, cur.Tjp, cur.w4);
        if (j > 1) fym = face_kw(a, cur.mjm, cur.Tjm, cur.T, cur.w4m);
        if (k < B.nk) fz = face_kw(a, m, cur.T, Tp, cur.w8);
      }
      if (lo_halo) fxm_own = face_kw(a, cur.mih, Tl, cur.T, cur.w0h);
    }
    R fxm = __shfl_up(fxp, 1, 64);
    if (lane == 0) fxm = fxm_own;
    if (live) {
      R Rv = 0;
      {
#pragma clang fp contract(off)
        if (i > 1) Rv = Rv + fxm;
        if (i < B.ni) Rv = Rv - fxp;
        if (j > 1) Rv = Rv + fym;
        if (j < B.nj) Rv = Rv - fyp;
        if (k > 1) Rv = Rv + fz_prev;
        if (k < B.nk) Rv = Rv - fz;
      }
      Rv = apply_bc_faces(a, B, i, j, k, Rv);
      const R dt = a.use_dtg ? (R)(*a.dtg) : cur.dt;
      int bad = 0;
      const R res = scaled_residual<PRIM>(a.tab, cur.m, cur.T, Rv, cur.vol, cur.qv, dt, &bad);
      if (IRSMODE) {
        stu(a.Rres, p, res);
      } else {
        const R Tn = new_state<PRIM>(a.tab, cur.m, cur.T, cur.TO, res, a.rkc, a.jam, &bad);
        stu(a.Tout, p, Tn);
        if (Tn != Tn || Tn < (R)0) flag_state(a.status, a.irk, t.b, i, j, k);
        const double d = (double)Tn - (double)cur.TO;
        rsq += d * d;
      }
      if (bad) flag_error(a.status, FUSS_ETABLE, t.b, i, j, k);
    }
    fz_prev = fz;
    cur = nxt;
    Tp = Tp2;
  }
  if (!IRSMODE && a.last) block_sum(rsq, a.partials, wg);
}
