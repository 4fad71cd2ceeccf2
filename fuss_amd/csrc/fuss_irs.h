// fuss_irs.h -- implicit residual smoothing kernels (Jacobi pairs and
// tridiagonal lines) with the fused RK update, gfx950.  Included by the
// IRS launch units.
#pragma once

#include "fuss_dev.h"

template <int DIR, int NDIR, bool UPDATE, bool PRIM, typename R>
__global__ __launch_bounds__(TILE_T) void k_irs(IrsArgs<R> a) {
  if (a.halt && *a.halt) return;
  const int wg = xcd_remap(blockIdx.x, gridDim.x);
  const Tile t = a.tiles[wg];
  const BlockInfo B = a.blk[t.b];
  const int i = t.i0 + (threadIdx.x % TILE_X), j = t.j0 + (threadIdx.x / TILE_X), k = t.k0;
  double rsq = 0.0;
  if (i <= B.ni && j <= B.nj) {
#pragma clang fp contract(off)
    const long p = B.base + (i - 1) + (long)(j - 1) * B.sy + (long)(k - 1) * B.sz;
    const long st = DIR == 0 ? 1 : (DIR == 1 ? (long)B.sy : B.sz);
    const int s = DIR == 0 ? i : (DIR == 1 ? j : k);
    const int n = DIR == 0 ? B.ni : (DIR == 1 ? B.nj : B.nk);
    const R* __restrict__ Rin = a.Rin;
    R glo = 0, ghi = 0;
    if (s <= 2) {
      const int f = 2 * DIR + 1;
      const int i1 = DIR == 0 ? 1 : i, j1 = DIR == 1 ? 1 : j, k1 = DIR == 2 ? 1 : k;
      const BCRec& e = a.bc[a.fmap[face_map_index<R>(B, f, i1, j1, k1)]];
      glo = e.type == 101 ? Rin[e.rdonor] : Rin[p - (long)(s - 1) * st];
    }
    if (s >= n - 1) {
      const int f = 2 * DIR + 2;
      const int i1 = DIR == 0 ? B.ni : i, j1 = DIR == 1 ? B.nj : j, k1 = DIR == 2 ? B.nk : k;
      const BCRec& e = a.bc[a.fmap[face_map_index<R>(B, f, i1, j1, k1)]];
      ghi = e.type == 101 ? Rin[e.rdonor] : Rin[p + (long)(n - s) * st];
    }
    auto r = [&](int d) -> R {
      const int q = s + d;
      if (q < 1) return glo;
      if (q > n) return ghi;
      return Rin[p + d * st];
    };
    const R beta = a.beta, den = a.den, inv = a.inv;
    // first sweep at s-1 and s+1 (ghost cells keep r*)
    const R rm = (s - 1 < 1) ? glo : irs_div(r(-1) + beta * (r(-2) + r(0)), den, inv);
    const R rp = (s + 1 > n) ? ghi : irs_div(r(1) + beta * (r(0) + r(2)), den, inv);
    const R rs = irs_div(r(0) + beta * (rm + rp), den, inv);
    if (!UPDATE) {
      a.Rout[p] = rs;
    } else {
      const R T = ((PRIM && a.jam) ? a.TO : a.T)[p], TO = a.TO[p];  // four-stage prim: T unused
      int bad = 0;
      const R Tn = new_state<PRIM>(a.tab, a.mat[p], T, TO, rs, a.rkc, a.jam, &bad);
      a.Tout[p] = Tn;
      if (Tn != Tn || Tn < (R)0) flag_state(a.status, a.irk, t.b, i, j, k);
      if (bad) flag_error(a.status, FUSS_ETABLE, t.b, i, j, k);
      const double d = (double)Tn - (double)TO;
      rsq = d * d;
    }
  }
  if (UPDATE && a.last) block_sum(rsq, a.partials, wg);
}

// IRS ghost of the line end on face f (Ghost_Residual_Extrapolation,
// Lib_IRS.f90:170-184): the donor's R for a connection, else the boundary
// cell's own R (zero gradient).
template <typename R>
__device__ __forceinline__ R irs_ghost(const IrsArgs<R>& a, const BlockInfo& B, int f, int i, int j, int k,
                                       R own) {
  if (!(B.conn & (1 << (f - 1)))) return own;  // no connection on this face
  const int e = a.fmap[face_map_index<R>(B, f, i, j, k)];
  const BCRec& r = a.bc[e];
  return r.type == 101 ? ldu(a.Rin, (unsigned)r.rdonor) : own;
}


// x then y smoothing of one plane tile of 64 x 16 cells (Residual_Smoothing
// directions 1 and 2, Lib_IRS.f90:25-38) through LDS: R is read once.  Used
// when no block has a connection on a j face (the y ghosts are then the
// x-smoothed boundary rows).  2-D levels finish the stage here (UPDATE).
template <int NDIR, bool UPDATE, bool PRIM, typename R>
__global__ __launch_bounds__(TILE_T) void k_irs_xy(IrsArgs<R> a) {
  if (a.halt && *a.halt) return;
  const int wg = xcd_remap(blockIdx.x, gridDim.x);
  const Tile t = a.tiles[wg];
  const BlockInfo B = a.blk[t.b];
  const int k = t.k0;
  __shared__ R rin[XY_RJ][XY_RI];
  __shared__ R xs[XY_RJ][TILE_X];
  const unsigned corner = (unsigned)(B.base + (t.i0 - 3) + (long)(t.j0 - 3) * B.sy + (long)(k - 1) * B.sz);
  for (int q = threadIdx.x; q < XY_RJ * XY_RI; q += TILE_T) {
    const int ii = q % XY_RI, jj = q / XY_RI;
    const int gi = t.i0 - 2 + ii, gj = t.j0 - 2 + jj;
    R v = 0;
    if (gi >= 1 && gi <= B.ni && gj >= 1 && gj <= B.nj) v = ldu(a.Rin, corner + ii + jj * B.sy);
    rin[jj][ii] = v;
  }
  __syncthreads();
  const R beta = a.beta, den = a.den;
  // x smoothing of rows j0-2 .. j0+17 inside the block
  for (int q = threadIdx.x; q < XY_RJ * TILE_X; q += TILE_T) {
    const int ii = q % TILE_X, jj = q / TILE_X;
    const int i = t.i0 + ii, j = t.j0 - 2 + jj;
    if (i > B.ni || j < 1 || j > B.nj) continue;
    R glo = 0, ghi = 0;
    if (i <= 2) glo = irs_ghost(a, B, 1, 1, j, k, rin[jj][ii + 2 - (i - 1)]);
    if (i >= B.ni - 1) ghi = irs_ghost(a, B, 2, B.ni, j, k, rin[jj][ii + 2 + (B.ni - i)]);
    xs[jj][ii] = jacobi2(rin[jj][ii], rin[jj][ii + 1], rin[jj][ii + 2], rin[jj][ii + 3], rin[jj][ii + 4], i, B.ni,
                         glo, ghi, beta, den, a.inv);
  }
  __syncthreads();
  // y smoothing of the 16 output rows; y ghosts are the x-smoothed rows 1, nj
  double rsq = 0.0;
  for (int q = threadIdx.x; q < XY_TJ * TILE_X; q += TILE_T) {
    const int ii = q % TILE_X, jr = q / TILE_X;
    const int i = t.i0 + ii, j = t.j0 + jr;
    if (i > B.ni || j > B.nj) continue;
    const int jj = jr + 2;  // LDS row of j
    const R glo = (j <= 2) ? xs[jj - (j - 1)][ii] : (R)0;
    const R ghi = (j >= B.nj - 1) ? xs[jj + (B.nj - j)][ii] : (R)0;
    const R rs = jacobi2(xs[jj - 2][ii], xs[jj - 1][ii], xs[jj][ii], xs[jj + 1][ii], xs[jj + 2][ii], j, B.nj, glo, ghi,
                         beta, den, a.inv);
    const unsigned p = (unsigned)(B.base + (i - 1) + (long)(j - 1) * B.sy + (long)(k - 1) * B.sz);
    if (!UPDATE) {
      stu(a.Rout, p, rs);
    } else {
      const R T = ldu((PRIM && a.jam) ? a.TO : a.T, p), TO = ldu(a.TO, p);  // four-stage prim: T unused
      int bad = 0;
      const int m = B.umat >= 0 ? B.umat : ldu(a.mat, p);
      const R Tn = new_state<PRIM>(a.tab, m, T, TO, rs, a.rkc, a.jam, &bad);
      stu(a.Tout, p, Tn);
      if (Tn != Tn || Tn < (R)0) flag_state(a.status, a.irk, t.b, i, j, k);
      if (bad) flag_error(a.status, FUSS_ETABLE, t.b, i, j, k);
      const double d = (double)Tn - (double)TO;
      rsq += d * d;
    }
  }
  if (UPDATE && a.last) block_sum(rsq, a.partials, wg);
}

template <typename R>
__device__ __forceinline__ R tri_ghost(const TriArgs<R>& a, const BlockInfo& B, int f, int i, int j, int k, R own) {
  if (!(B.conn & (1 << (f - 1)))) return own;
  const BCRec& r = a.bc[a.fmap[face_map_index<R>(B, f, i, j, k)]];
  return r.type == 101 ? ldu(a.Rin, (unsigned)r.rdonor) : own;
}

// y (DIR 1) and z (DIR 2) lines: a 64x4 tile of line origins, i fastest, so
// every step of the sweep is a coalesced row access.
template <int DIR, typename R>
__global__ __launch_bounds__(TILE_T) void k_tri_yz(TriArgs<R> a) {
  if (a.halt && *a.halt) return;
  const Tile t = a.tiles[blockIdx.x];
  const BlockInfo B = a.blk[t.b];
  const int i = t.i0 + threadIdx.x % TILE_X, o2 = threadIdx.x / TILE_X;
  // DIR 1: lines (i, k = k0 + o2); DIR 2: lines (i, j = j0 + o2)
  const int j = DIR == 1 ? 1 : t.j0 + o2, k = DIR == 1 ? t.k0 + o2 : 1;
  const int n = DIR == 1 ? B.nj : B.nk;
  if (i > B.ni || (DIR == 1 ? k > B.nk : j > B.nj)) return;
  const unsigned st = DIR == 1 ? (unsigned)B.sy : (unsigned)B.sz;
  const unsigned p1 = (unsigned)(B.base + (i - 1) + (long)(j - 1) * B.sy + (long)(k - 1) * B.sz);
  const double2* tab = a.tri + a.trioff[3 * t.b + DIR];
  const R beta = a.beta;
  const R r1 = ldu(a.Rin, p1), rn = ldu(a.Rin, p1 + (unsigned)(n - 1) * st);
  const R glo = DIR == 1 ? tri_ghost(a, B, 3, i, 1, k, r1) : tri_ghost(a, B, 5, i, j, 1, r1);
  const R ghi = DIR == 1 ? tri_ghost(a, B, 4, i, B.nj, k, rn) : tri_ghost(a, B, 6, i, j, B.nk, rn);
  R d = 0;
  unsigned p = p1;
  for (int s = 1; s <= n; s++, p += st) {
#pragma clang fp contract(off)
    R r = ldu(a.Rin, p);
    if (s == 1) r = r + beta * glo;
    if (s == n) r = r + beta * ghi;
    const R den = (R)tab[s - 1].x;
    d = s == 1 ? r / den : (r + beta * d) / den;
    stu(a.Rout, p, d);
  }
  R x = d;
  p = p1 + (unsigned)(n - 1) * st;
  for (int s = n - 1; s >= 1; s--) {
#pragma clang fp contract(off)
    p -= st;
    x = ldu((const R*)a.Rout, p) - (R)tab[s - 1].y * x;
    stu(a.Rout, p, x);
  }
}

// x lines: one wave per 64 consecutive rows j of a plane; the lines pass
// through LDS in chunks of 16 cells so that global accesses are whole
// 128-byte row segments.
constexpr int TRI_CH = 16;

template <typename R>
__global__ __launch_bounds__(64) void k_tri_x(TriArgs<R> a) {
  if (a.halt && *a.halt) return;
  const Tile t = a.tiles[blockIdx.x];
  const BlockInfo B = a.blk[t.b];
  const int l = threadIdx.x;                // this lane's line: row j0 + l
  const int j = t.j0 + l, k = t.k0, n = B.ni;
  const bool live = j <= B.nj;
  __shared__ R ch[TRI_CH][TILE_X + 1];
  const unsigned row0 = (unsigned)(B.base + (long)(t.j0 - 1) * B.sy + (long)(k - 1) * B.sz);  // (i=1, j0, k)
  const double2* tab = a.tri + a.trioff[3 * t.b];
  const R beta = a.beta;
  R glo = 0, ghi = 0;
  if (live) {
    const unsigned pl = row0 + (unsigned)l * (unsigned)B.sy;
    glo = tri_ghost(a, B, 1, 1, j, k, ldu(a.Rin, pl));
    ghi = tri_ghost(a, B, 2, B.ni, j, k, ldu(a.Rin, pl + (unsigned)(n - 1)));
  }
  const int nch = (n + TRI_CH - 1) / TRI_CH;
  // chunk c holds cells i = c*16+1 .. c*16+16 of the 64 rows: element e of a
  // lane's 16 loads is row e/16, cell e%16 (16 lanes per 128-byte segment)
  auto chunk_io = [&](int c, bool load, const R* src, R* dst) {
    for (int q = 0; q < TRI_CH; q++) {
      const int e = l + 64 * q, rr = e / TRI_CH, ii = e % TRI_CH;
      const int i = c * TRI_CH + ii + 1;
      if (t.j0 + rr <= B.nj && i <= n) {
        const unsigned p = row0 + (unsigned)rr * (unsigned)B.sy + (unsigned)(i - 1);
        if (load) ch[ii][rr] = ldu(src, p);
        else stu(dst, p, ch[ii][rr]);
      }
    }
  };
  R d = 0;
  for (int c = 0; c < nch; c++) {
    chunk_io(c, true, a.Rin, nullptr);
    __syncthreads();
    if (live) {
      for (int ii = 0; ii < TRI_CH; ii++) {
#pragma clang fp contract(off)
        const int s = c * TRI_CH + ii + 1;
        if (s > n) break;
        R r = ch[ii][l];
        if (s == 1) r = r + beta * glo;
        if (s == n) r = r + beta * ghi;
        const R den = (R)tab[s - 1].x;
        d = s == 1 ? r / den : (r + beta * d) / den;
        ch[ii][l] = d;
      }
    }
    __syncthreads();
    chunk_io(c, false, nullptr, a.Rout);
    __syncthreads();
  }
  __threadfence_block();
  R x = d;
  for (int c = nch - 1; c >= 0; c--) {
    chunk_io(c, true, a.Rout, nullptr);
    __syncthreads();
    if (live) {
      for (int ii = TRI_CH - 1; ii >= 0; ii--) {
#pragma clang fp contract(off)
        const int s = c * TRI_CH + ii + 1;
        if (s >= n) continue;  // x_n = d_n stays
        x = ch[ii][l] - (R)tab[s - 1].y * x;
        ch[ii][l] = x;
      }
    }
    __syncthreads();
    chunk_io(c, false, nullptr, a.Rout);
    __syncthreads();
  }
}

// The RK update from the smoothed residual (after the last tridiagonal
// direction), with the residual partials of the step.
template <bool PRIM, typename R>
__global__ __launch_bounds__(TILE_T) void k_rk_update(IrsArgs<R> a) {
  if (a.halt && *a.halt) return;
  const int wg = xcd_remap(blockIdx.x, gridDim.x);
  const Tile t = a.tiles[wg];
  const BlockInfo B = a.blk[t.b];
  const int i = t.i0 + (threadIdx.x % TILE_X), j = t.j0 + (threadIdx.x / TILE_X), k = t.k0;
  double rsq = 0.0;
  if (i <= B.ni && j <= B.nj) {
    const unsigned p = (unsigned)(B.base + (i - 1) + (long)(j - 1) * B.sy + (long)(k - 1) * B.sz);
    const R T = ldu((PRIM && a.jam) ? a.TO : a.T, p), TO = ldu(a.TO, p), rs = ldu(a.Rin, p);
    int bad = 0;
    const int m = B.umat >= 0 ? B.umat : ldu(a.mat, p);
    const R Tn = new_state<PRIM>(a.tab, m, T, TO, rs, a.rkc, a.jam, &bad);
    stu(a.Tout, p, Tn);
    if (Tn != Tn || Tn < (R)0) flag_state(a.status, a.irk, t.b, i, j, k);
    if (bad) flag_error(a.status, FUSS_ETABLE, t.b, i, j, k);
    const double dd = (double)Tn - (double)TO;
    rsq = dd * dd;
  }
  if (a.last) block_sum(rsq, a.partials, wg);
}

// z smoothing (direction 3) fused with the RK update: a 64x4 column marches
// through KC planes with r(k-2..k+2) in registers, so R is read once.
template <bool PRIM, typename R, int KCT>
__global__ __launch_bounds__(TILE_T) void k_irs_z(IrsArgs<R> a) {
  if (a.halt && *a.halt) return;
  const int wg = xcd_remap(blockIdx.x, gridDim.x);
  const Tile t = a.tiles[wg];
  const BlockInfo B = a.blk[t.b];
  const int i = t.i0 + (threadIdx.x % TILE_X), j = t.j0 + (threadIdx.x / TILE_X);
  const int k0 = t.k0, k1 = min(B.nk, k0 + KCT - 1), n = B.nk;
  double rsq = 0.0;
  if (i <= B.ni && j <= B.nj) {
    const unsigned sz = (unsigned)B.sz;
    const unsigned p1 = (unsigned)(B.base + (i - 1) + (long)(j - 1) * B.sy);  // plane k = 1
    const R glo = (k0 <= 2) ? irs_ghost(a, B, 5, i, j, 1, ldu(a.Rin, p1)) : (R)0;
    const R ghi = (k1 >= n - 1) ? irs_ghost(a, B, 6, i, j, n, ldu(a.Rin, p1 + (n - 1) * sz)) : (R)0;
    auto rd = [&](int kk) -> R { return (kk >= 1 && kk <= n) ? ldu(a.Rin, p1 + (kk - 1) * sz) : (R)0; };
    R w0 = rd(k0 - 2), w1 = rd(k0 - 1), w2 = rd(k0), w3 = rd(k0 + 1), w4 = rd(k0 + 2);
    const int m = B.umat >= 0 ? B.umat : -1;
    for (int k = k0; k <= k1; k++) {
      const R rs = jacobi2(w0, w1, w2, w3, w4, k, n, glo, ghi, a.beta, a.den, a.inv);
      const unsigned p = p1 + (k - 1) * sz;
      const R T = ldu((PRIM && a.jam) ? a.TO : a.T, p), TO = ldu(a.TO, p);  // four-stage prim: T unused
      int bad = 0;
      const R Tn = new_state<PRIM>(a.tab, m >= 0 ? m : ldu(a.mat, p), T, TO, rs, a.rkc, a.jam, &bad);
      stu(a.Tout, p, Tn);
      if (Tn != Tn || Tn < (R)0) flag_state(a.status, a.irk, t.b, i, j, k);
      if (bad) flag_error(a.status, FUSS_ETABLE, t.b, i, j, k);
      const double d = (double)Tn - (double)TO;
      rsq += d * d;
      w0 = w1; w1 = w2; w2 = w3; w3 = w4;
      w4 = rd(k + 3);
    }
  }
  if (a.last) block_sum(rsq, a.partials, wg);
}

// All three directions of Residual_Smoothing (Lib_IRS.f90:14-40) and the RK
// update in one pass, for 3-D levels without connections on j or k faces: a
// 64xTY column marches through its KC planes.  Each R plane enters LDS once
// with a 2-cell i/j halo, is smoothed along x then y there (the y ghosts are
// the x-smoothed boundary rows, as in k_irs_xy), and its 64xTY result joins a
// register window s(q-4..q) that the z sweeps read (z ghosts: the
// xy-smoothed boundary planes).  R is read from HBM once per stage instead of
// three times (x-y pass in, z pass in) plus once written.  Planes are
// fetched two ahead into two register sets that alternate over a 2-plane
// unrolled loop; every load is unconditional (clamped plane, valid cells for
// out-of-block lanes) so the waits leave the later planes in flight.
// XCON: some block has a connection on an i face (x ghosts from the donors).
// TY: rows of the column (4 or 8; one wave per row): the 8-row column reads
// 12 instead of 16 halo rows per 8 rows of output and sweeps 12 x lines.
template <int TY>
struct Xyz {
  static constexpr int TT = TILE_X * TY;                // threads
  static constexpr int RJ = TY + 4, RI = TILE_X + 4;    // halo box
  static constexpr int NLD = (RJ * RI + TT - 1) / TT;   // plane loads per thread
};

template <typename R, int NLD>
struct XyzPlane {
  R v[NLD];
};

// Element u of a thread's share of the tile's (TY+4) x (TILE_X+4) halo box
// lies in the block.
template <int TY>
__device__ __forceinline__ bool xyz_in(const BlockInfo& B, int i0, int j0, int u) {
  using G = Xyz<TY>;
  const int e = threadIdx.x + u * G::TT;
  const int gi = i0 - 2 + e % G::RI, gj = j0 - 2 + e / G::RI;
  return e < G::RJ * G::RI && gi >= 1 && gi <= B.ni && gj >= 1 && gj <= B.nj;
}

// The R values of plane q in the halo box, raw: cells outside the block read
// the box's first in-block cell, and the LDS store writes 0 for them (the
// mask is applied there, so no wait for these loads is placed before).
template <int TY, typename R>
__device__ __forceinline__ void xyz_fetch(XyzPlane<R, Xyz<TY>::NLD>& r, const R* __restrict__ Rin,
                                          const BlockInfo& B, int i0, int j0, int q) {
  using G = Xyz<TY>;
  const unsigned corner = (unsigned)(B.base + (i0 - 3) + (long)(j0 - 3) * B.sy + (long)(q - 1) * B.sz);
  // first in-block cell of the box: (max(i0-2,1), max(j0-2,1))
  const unsigned safe = corner + (i0 > 2 ? 0u : (unsigned)(3 - i0)) + (j0 > 2 ? 0u : (unsigned)(3 - j0)) * (unsigned)B.sy;
#pragma unroll
  for (int u = 0; u < G::NLD; u++) {
    const int e = threadIdx.x + u * G::TT;
    const int ii = e % G::RI, jj = e / G::RI;
    r.v[u] = ldu(Rin, xyz_in<TY>(B, i0, j0, u) ? corner + ii + jj * (unsigned)B.sy : safe);
  }
}

template <bool PRIM, bool XCON, bool P2, bool DT, typename R, int KCT, int TY>
__global__ __launch_bounds__(Xyz<TY>::TT) void k_irs_xyz(IrsArgs<R> a) {
  using G = Xyz<TY>;
  if (a.halt && *a.halt) return;
  const int wg = xcd_remap(blockIdx.x, gridDim.x);
  const Tile t = a.tiles[wg];
  const BlockInfo B = a.blk[t.b];
  const int tx = threadIdx.x % TILE_X, ty = threadIdx.x / TILE_X;
  const int i = t.i0 + tx, j = t.j0 + ty;
  const int n = B.nk, k0 = t.k0, k1 = min(n, k0 + KCT - 1);
  const bool live = i <= B.ni && j <= B.nj;
  __shared__ R rin[G::RJ][G::RI];
  __shared__ R xs[G::RJ][TILE_X];
  const R beta = a.beta, den = a.den, inv = a.inv;
  // window w[0..4] = s(q-4 .. q); planes outside 1..n are 0 (jacobi2 selects
  // the ghosts there)
  R w0 = 0, w1 = 0, w2 = 0, w3 = 0, w4 = 0;
  R zlo = 0, zhi = 0;  // xy-smoothed planes 1 and n (z ghosts of walls)
  const int qa = max(1, k0 - 2), qb = min(n, k1 + 2), qe = k1 + 2;
  const int ci = live ? i : 1, cj = live ? j : 1;
  const unsigned p1 = (unsigned)(B.base + (ci - 1) + (long)(cj - 1) * B.sy);  // plane 1 of the lane's column
  const unsigned sz = (unsigned)B.sz;
  double rsq = 0.0, dtmin = 1e300;
  // next-step dt on a rectilinear level: sum 1/dl^2 = (s_x + s_y) + s_z from
  // the per-axis factors (bit-identical to invL2), the column's s_z in LDS
  __shared__ double qzs[KCT];
  const bool rdt = DT && a.rect;
  double qxy = 0.0;
  if (rdt) {
    const RectAxes rx_ = rect_axes(a.rect, B);
    if (threadIdx.x < KCT) qzs[threadIdx.x] = rx_.qz[min(k0 + (int)threadIdx.x, n)];
    qxy = rx_.qx[ci] + rx_.qy[cj];
  }
  bool inb[G::NLD];
#pragma unroll
  for (int u = 0; u < G::NLD; u++) inb[u] = xyz_in<TY>(B, t.i0, t.j0, u);
  // the tile's x lines and the wave's y line (one row j per wave) away from
  // the block faces: no ghost selection
  const bool xin = t.i0 >= 3 && t.i0 + TILE_X - 1 <= B.ni - 2;
  const int ju = t.j0 + __builtin_amdgcn_readfirstlane(ty);
  const bool yin = ju >= 3 && ju <= B.nj - 2;
  XyzPlane<R, G::NLD> pa, pb;
  xyz_fetch<TY>(pa, a.Rin, B, t.i0, t.j0, qa);
  xyz_fetch<TY>(pb, a.Rin, B, t.i0, t.j0, min(qa + 1, qb));
  // one step: plane q (staged in `pre`) is smoothed in x-y (if q <= qb), the
  // window shifts, and plane q-2 is output (if it is in the column)
  auto step = [&](XyzPlane<R, G::NLD>& pre, int q) {
    const int ko = q - 2;
    const bool out = ko >= k0;
    const unsigned po = p1 + (unsigned)(max(ko, 1) - 1) * sz;
    // the four-stage update in primitive variables (new = old + a_k res) does
    // not read the stage state: its load then hits T^n's line (no HBM bytes)
    const R T = ldn((PRIM && a.jam) ? a.TO : a.T, po), TO = ldn(a.TO, po);
    int m = B.umat;
    if (m < 0) m = ldu(a.mat, po);
    R s = 0;
    if (q <= qb) {
#pragma unroll
      for (int u = 0; u < G::NLD; u++) {
        const int e = threadIdx.x + u * G::TT;
        if (e < G::RJ * G::RI) (&rin[0][0])[e] = inb[u] ? pre.v[u] : (R)0;
      }
    }
    // plane q+2 flies while this one is smoothed (issued on every step, so
    // the wait for T and TO below can leave it outstanding)
    xyz_fetch<TY>(pre, a.Rin, B, t.i0, t.j0, min(q + 2, qb));
    if (q <= qb) {
      __syncthreads();
      // x sweeps of rows j0-2 .. j0+TY+1 (wave ty: rows ty, ty+TY, ...); a
      // tile whose cells are all >= 2 cells from the i faces selects no ghost
      // (uniform branch)
#pragma unroll
      for (int h = 0; h < (G::RJ + TY - 1) / TY; h++) {
        const int jj = ty + h * TY, jr = t.j0 - 2 + jj;
        if (jj < G::RJ && i <= B.ni && jr >= 1 && jr <= B.nj) {
          if (xin) {
            xs[jj][tx] = jacobi2_in<P2>(rin[jj][tx], rin[jj][tx + 1], rin[jj][tx + 2], rin[jj][tx + 3],
                                        rin[jj][tx + 4], beta, den, inv);
          } else {
            R glo = 0, ghi = 0;
            if (i <= 2) {
              glo = rin[jj][tx + 2 - (i - 1)];
              if (XCON) glo = irs_ghost(a, B, 1, 1, jr, q, glo);
            }
            if (i >= B.ni - 1) {
              ghi = rin[jj][tx + 2 + (B.ni - i)];
              if (XCON) ghi = irs_ghost(a, B, 2, B.ni, jr, q, ghi);
            }
            xs[jj][tx] = jacobi2_p<P2>(rin[jj][tx], rin[jj][tx + 1], rin[jj][tx + 2], rin[jj][tx + 3],
                                       rin[jj][tx + 4], i, B.ni, glo, ghi, beta, den, inv);
          }
        }
      }
      __syncthreads();
      // y sweeps of the lane's cell: the next plane's LDS store follows the
      // second barrier (rin no longer read), its x sweeps the next first
      // barrier (xs no longer read)
      if (live) {
        const int jj = ty + 2;
        if (yin) {
          s = jacobi2_in<P2>(xs[jj - 2][tx], xs[jj - 1][tx], xs[jj][tx], xs[jj + 1][tx], xs[jj + 2][tx], beta, den,
                             inv);
        } else {
          const R glo = (j <= 2) ? xs[jj - (j - 1)][tx] : (R)0;
          const R ghi = (j >= B.nj - 1) ? xs[jj + (B.nj - j)][tx] : (R)0;
          s = jacobi2_p<P2>(xs[jj - 2][tx], xs[jj - 1][tx], xs[jj][tx], xs[jj + 1][tx], xs[jj + 2][tx], j, B.nj,
                            glo, ghi, beta, den, inv);
        }
      }
      if (q == 1) zlo = s;
      if (q == n) zhi = s;
    }
    w0 = w1; w1 = w2; w2 = w3; w3 = w4; w4 = s;
    if (out && live) {
      const R rs = (ko >= 3 && ko <= n - 2) ? jacobi2_in<P2>(w0, w1, w2, w3, w4, beta, den, inv)
                                            : jacobi2_p<P2>(w0, w1, w2, w3, w4, ko, n, zlo, zhi, beta, den, inv);
      int bad = 0;
      const R Tn = new_state<PRIM>(a.tab, m, T, TO, rs, a.rkc, a.jam, &bad);
      stn(a.Tout, po, Tn);
      if (Tn != Tn || Tn < (R)0) flag_state(a.status, a.irk, t.b, i, j, ko);
      if (bad) flag_error(a.status, FUSS_ETABLE, t.b, i, j, ko);
      const double d = (double)Tn - (double)TO;
      rsq += d * d;
      if (DT) {
        // Compute_dt (Mod_dt.f90:26-84) at T^{n+1}, without the vnn ramp
        // (the host fuses only when it cannot apply); a separate instance of
        // the kernel, so the other stages keep the registers it needs
        int bad3 = 0;
        using V2 = typename Vec2<R>::T;
        using V4 = typename Vec4<R>::T;
        // (v, dv) rows: the interpolation v + (T - Tm) dv of tab_interp
        const int rw = tab_row(Tn, a.tab.Ti1, a.tab.Ti2, &bad3) - a.tab.Ti1;
        const unsigned roff = (unsigned)((long)m * a.tab.nT + rw);
        const V2 kk = ldu(reinterpret_cast<const V2*>(a.tab.kap2), roff);
        const V4 pp = ldu(reinterpret_cast<const V4*>(a.tab.pr4), roff);
        double dd;
        {
#pragma clang fp contract(off)
          const R fr = Tn - (R)(rw + a.tab.Ti1);
          const R kap = kk.x + fr * kk.y, rho = pp.x + fr * pp.y, cp = pp.z + fr * pp.w;
          const double alpha = (double)kap / (double)rho / (double)cp;
          dd = 0.5 * a.vnn / (alpha * (rdt ? qxy + qzs[ko - k0] : ldn(a.invL2, po)));
          dd = dd < 1e8 ? dd : 1e8;
        }
        stn(a.dt, po, (R)dd);
        dtmin = fmin(dtmin, dd);
        if (bad3) flag_error(a.status, FUSS_ETABLE, t.b, i, j, ko);
      }
    }
  };
  for (int q = qa; q <= qe; q += 2) {
    step(pa, q);
    if (q + 1 <= qe) step(pb, q + 1);
  }
  if (a.last) block_sum<G::TT>(rsq, a.partials, wg);
  if (DT) block_min<G::TT>(dtmin, a.pmin, wg);
}
