// fuss_refine.h -- steady-state defect correction on gfx950.
//
// The steady problem of the explicit solver is F(T) = 0 with, per cell,
//   b(T) = -R(T) + qvol V          (W; R from Fluxes + BC_Fluxes)
// (Mod_Fluxes.f90:10-126, Mod_BC_Fluxes.f90:10-86, Lib_Newstate.f90:241-256).
// One outer iteration of the refinement evaluates b in FP64 on the finest
// level and solves the correction equation
//   A d = b,   (A d)_c = sum_f c_f (d_c - d_nb) + D_c d_c
// approximately in the inner precision S, then T += d in FP64.  c_f is the
// conductance of interior face f at the current T (kappa of the left cell at
// the face-average T times the normal face weight, the terms of
// Diffusive_Flux that act on the normal difference, Lib_Diffusive.f90:10-45),
// D_c the derivative of the cell's BC fluxes with respect to its own T with
// kappa frozen.  Because the outer residual is exact FP64 and every inner
// term is proportional to d, the inner rounding error shrinks with d and the
// outer iteration converges to the FP64 fixed point.
//
// Inner solver: cell-centred multigrid over 2:1 aggregates (a direction stops
// coarsening at 2 cells), coarse conductances = (sum of the fine face
// conductances across the coarse face) / (aggregate centre distance in fine
// cells), D summed over children; restriction sums residuals, prolongation is
// the trilinear 27/9/3/1 interpolation of coarse2fine_prim
// (Lib_Multigrid.f90:128-220) on the correction.  Smoother: the multistage
// (Runge-Kutta form) sweep d_k = d_0 + a_k s(d_{k-1}), s = (b - A d)/diag(A),
// optionally followed by IRS line sweeps (Lib_IRS.f90:63-89), with a_k from
// the Chebyshev polynomial that damps the Jacobi-scaled spectrum on
// [lam_lo, lam_hi].  Blocks are independent (no connection faces).
#pragma once

namespace fuss {

template <typename S>
struct COp {
  const CBlk* blk;
  int nb, ncell;
  const S* cx;   // conductance of the cell's low i face (0 on the block boundary)
  const S* cy;
  const S* cz;
  const S* D;    // BC diagonal
  const S* b;    // right-hand side (W)
};

__device__ __forceinline__ int cmg_find_block(const CBlk* __restrict__ blk, int nb, int g) {
  int lo = 0, hi = nb - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (blk[mid].off <= g) lo = mid;
    else hi = mid - 1;
  }
  return lo;
}

// (A d)_g and diag(A)_g of cell (i,j,k) (0-based) of block B.
template <typename S>
__device__ __forceinline__ S cmg_Ad(const COp<S>& o, const S* __restrict__ d, const CBlk& B, int g, int i, int j,
                                    int k, S* diag) {
  const int sy = B.n[0], sz = B.n[0] * B.n[1];
  const S dc = d[g];
  S sc = o.D[g];
  S a = sc * dc;
  if (i > 0) { const S c = o.cx[g]; a += c * (dc - d[g - 1]); sc += c; }
  if (i < B.n[0] - 1) { const S c = o.cx[g + 1]; a += c * (dc - d[g + 1]); sc += c; }
  if (j > 0) { const S c = o.cy[g]; a += c * (dc - d[g - sy]); sc += c; }
  if (j < B.n[1] - 1) { const S c = o.cy[g + sy]; a += c * (dc - d[g + sy]); sc += c; }
  if (k > 0) { const S c = o.cz[g]; a += c * (dc - d[g - sz]); sc += c; }
  if (k < B.n[2] - 1) { const S c = o.cz[g + sz]; a += c * (dc - d[g + sz]); sc += c; }
  *diag = sc;
  return a;
}

struct CCell {
  int b, i, j, k;
};
__device__ __forceinline__ CCell cmg_cell(const CBlk& B, int bi, int g) {
  const int l = g - B.off;
  const int a = l / B.n[0];
  return CCell{bi, l - a * B.n[0], a % B.n[1], a / B.n[1]};
}

// One smoother stage.  WRITE_S: store s = (b - A d)/diag for the IRS passes;
// else the stage update out = d0 + alpha s.
template <typename S, bool WRITE_S>
__global__ __launch_bounds__(256) void k_cmg_stage(COp<S> o, const S* __restrict__ din, const S* __restrict__ d0,
                                                   S* __restrict__ out, S alpha) {
  const int g = xcd_remap(blockIdx.x, gridDim.x) * 256 + threadIdx.x;
  if (g >= o.ncell) return;
  const int bi = cmg_find_block(o.blk, o.nb, g);
  const CBlk B = o.blk[bi];
  const CCell q = cmg_cell(B, bi, g);
  S dg;
  const S Ad = cmg_Ad(o, din, B, g, q.i, q.j, q.k, &dg);
  const S s = dg > (S)0 ? (o.b[g] - Ad) / dg : (S)0;
  if (WRITE_S) out[g] = s;
  else out[g] = d0[g] + alpha * s;
}

// IRS along DIR on the smoother increment: two Jacobi sweeps with
// zero-gradient line ends; UPDATE applies the stage update.
template <int DIR, bool UPDATE, typename S>
__global__ __launch_bounds__(256) void k_cmg_irs(COp<S> o, const S* __restrict__ in, S* __restrict__ out, S beta,
                                                 S den, S inv, const S* __restrict__ d0, S alpha) {
  const int g = xcd_remap(blockIdx.x, gridDim.x) * 256 + threadIdx.x;
  if (g >= o.ncell) return;
  const int bi = cmg_find_block(o.blk, o.nb, g);
  const CBlk B = o.blk[bi];
  const CCell q = cmg_cell(B, bi, g);
  const int n = B.n[DIR];
  const int s = DIR == 0 ? q.i : (DIR == 1 ? q.j : q.k);
  const int st = DIR == 0 ? 1 : (DIR == 1 ? B.n[0] : B.n[0] * B.n[1]);
  auto r = [&](int dd) -> S {
    int t = s + dd;
    t = t < 0 ? 0 : (t > n - 1 ? n - 1 : t);
    return in[g + (t - s) * st];
  };
  const S glo = in[g - s * st], ghi = in[g + (n - 1 - s) * st];
  const S v = jacobi2(r(-2), r(-1), r(0), r(1), r(2), s + 1, n, glo, ghi, beta, den, inv);
  if (UPDATE) out[g] = d0[g] + alpha * v;
  else out[g] = v;
}

// Coarse right-hand side: the sum of the children's residuals b - A d
// (accumulated in FP64); the coarse correction starts from zero.
template <typename S>
__global__ __launch_bounds__(256) void k_cmg_restrict(COp<S> f, const S* __restrict__ df, COp<S> c, S* __restrict__ bc,
                                                      S* __restrict__ dc) {
  const int g = xcd_remap(blockIdx.x, gridDim.x) * 256 + threadIdx.x;
  if (g >= c.ncell) return;
  const int bi = cmg_find_block(c.blk, c.nb, g);
  const CBlk C = c.blk[bi];
  const CBlk F = f.blk[bi];
  const CCell q = cmg_cell(C, bi, g);
  double acc = 0.0;
  const int i0 = C.r[0] * q.i, j0 = C.r[1] * q.j, k0 = C.r[2] * q.k;
  const int i1 = min(i0 + C.r[0], F.n[0]), j1 = min(j0 + C.r[1], F.n[1]), k1 = min(k0 + C.r[2], F.n[2]);
  for (int kk = k0; kk < k1; kk++)
    for (int jj = j0; jj < j1; jj++)
      for (int ii = i0; ii < i1; ii++) {
        const int gf = F.off + ii + F.n[0] * (jj + F.n[1] * kk);
        S dg;
        const S Ad = cmg_Ad(f, df, F, gf, ii, jj, kk, &dg);
        acc += (double)(f.b[gf] - Ad);
      }
  bc[g] = (S)acc;
  dc[g] = (S)0;
}

// d_fine += P d_coarse: per direction weights 3/4 (own aggregate) and 1/4
// (the neighbour on the cell's side, index clamped into the block).
template <typename S>
__global__ __launch_bounds__(256) void k_cmg_prolong(COp<S> f, S* __restrict__ df, COp<S> c,
                                                     const S* __restrict__ dc) {
  const int g = xcd_remap(blockIdx.x, gridDim.x) * 256 + threadIdx.x;
  if (g >= f.ncell) return;
  const int bi = cmg_find_block(f.blk, f.nb, g);
  const CBlk F = f.blk[bi];
  const CBlk C = c.blk[bi];
  const CCell q = cmg_cell(F, bi, g);
  const int fi[3] = {q.i, q.j, q.k};
  int I[3][2];
  S w[3][2];
  int nw[3];
  for (int d = 0; d < 3; d++) {
    const int own = fi[d] / C.r[d];
    I[d][0] = own;
    if (C.r[d] == 2) {
      int nb = (fi[d] & 1) ? own + 1 : own - 1;
      nb = nb < 0 ? 0 : (nb > C.n[d] - 1 ? C.n[d] - 1 : nb);
      I[d][1] = nb;
      w[d][0] = (S)0.75;
      w[d][1] = (S)0.25;
      nw[d] = 2;
    } else {
      I[d][1] = own;
      w[d][0] = (S)1;
      w[d][1] = (S)0;
      nw[d] = 1;
    }
  }
  S acc = 0;
  for (int a = 0; a < nw[2]; a++)
    for (int bb = 0; bb < nw[1]; bb++)
      for (int e = 0; e < nw[0]; e++)
        acc += w[2][a] * w[1][bb] * w[0][e] * dc[C.off + I[0][e] + C.n[0] * (I[1][bb] + C.n[1] * I[2][a])];
  df[g] = df[g] + acc;
}

// Coarse operator from the finer one.
template <typename S>
__global__ __launch_bounds__(256) void k_cmg_coarsen(COp<S> f, const CBlk* __restrict__ cblk, int nb, int ncell,
                                                     S* __restrict__ cx, S* __restrict__ cy, S* __restrict__ cz,
                                                     S* __restrict__ D) {
  const int g = xcd_remap(blockIdx.x, gridDim.x) * 256 + threadIdx.x;
  if (g >= ncell) return;
  const int bi = cmg_find_block(cblk, nb, g);
  const CBlk C = cblk[bi];
  const CBlk F = f.blk[bi];
  const CCell q = cmg_cell(C, bi, g);
  const int ci[3] = {q.i, q.j, q.k};
  int lo[3], hi[3];
  for (int d = 0; d < 3; d++) {
    lo[d] = C.r[d] * ci[d];
    hi[d] = min(lo[d] + C.r[d], F.n[d]);
  }
  auto fidx = [&](int ii, int jj, int kk) { return F.off + ii + F.n[0] * (jj + F.n[1] * kk); };
  double sd = 0.0;
  for (int kk = lo[2]; kk < hi[2]; kk++)
    for (int jj = lo[1]; jj < hi[1]; jj++)
      for (int ii = lo[0]; ii < hi[0]; ii++) sd += (double)f.D[fidx(ii, jj, kk)];
  D[g] = (S)sd;
  // low-face conductance in direction d: fine faces at fine index lo[d]
  const S* fc[3] = {f.cx, f.cy, f.cz};
  S* oc[3] = {cx, cy, cz};
  for (int d = 0; d < 3; d++) {
    double s = 0.0;
    if (ci[d] > 0) {
      int a0[3] = {lo[0], lo[1], lo[2]}, a1[3] = {hi[0], hi[1], hi[2]};
      a1[d] = a0[d] + 1;
      for (int kk = a0[2]; kk < a1[2]; kk++)
        for (int jj = a0[1]; jj < a1[1]; jj++)
          for (int ii = a0[0]; ii < a1[0]; ii++) s += (double)fc[d][fidx(ii, jj, kk)];
      // centre distance of the two aggregates in fine cells: 2, or 1.5 when
      // this aggregate is a single trailing cell; 1 without coarsening
      const double dist = C.r[d] == 1 ? 1.0 : (hi[d] - lo[d] == 2 ? 2.0 : 1.5);
      s /= dist;
    }
    oc[d][g] = (S)s;
  }
}

// Exact solve of a coarsest block of at most CMG_DIRECT cells: one work-item
// per block, Gaussian elimination in FP64 on the SPD system.
constexpr int CMG_DIRECT = 8;
template <typename S>
__global__ void k_cmg_direct(COp<S> o, S* __restrict__ d) {
  const int bi = blockIdx.x * blockDim.x + threadIdx.x;
  if (bi >= o.nb) return;
  const CBlk B = o.blk[bi];
  const int n = B.n[0] * B.n[1] * B.n[2];
  if (n > CMG_DIRECT) return;
  double A[CMG_DIRECT][CMG_DIRECT + 1];
  for (int r = 0; r < n; r++)
    for (int cc = 0; cc <= n; cc++) A[r][cc] = 0.0;
  const int sy = B.n[0], sz = B.n[0] * B.n[1];
  for (int r = 0; r < n; r++) {
    const int g = B.off + r;
    const int i = r % B.n[0], j = (r / B.n[0]) % B.n[1], k = r / sz;
    A[r][r] += (double)o.D[g];
    A[r][n] = (double)o.b[g];
    auto link = [&](double cval, int r2) {
      A[r][r] += cval;
      A[r][r2] -= cval;
    };
    if (i > 0) link((double)o.cx[g], r - 1);
    if (i < B.n[0] - 1) link((double)o.cx[g + 1], r + 1);
    if (j > 0) link((double)o.cy[g], r - sy);
    if (j < B.n[1] - 1) link((double)o.cy[g + sy], r + sy);
    if (k > 0) link((double)o.cz[g], r - sz);
    if (k < B.n[2] - 1) link((double)o.cz[g + sz], r + sz);
  }
  for (int p = 0; p < n; p++) {
    const double piv = A[p][p];
    if (!(piv > 0.0)) continue;  // singular (no BC diagonal reaches this block): leave d = 0
    for (int r = p + 1; r < n; r++) {
      const double m = A[r][p] / piv;
      for (int cc = p; cc <= n; cc++) A[r][cc] -= m * A[p][cc];
    }
  }
  double x[CMG_DIRECT];
  for (int p = n - 1; p >= 0; p--) {
    double s = A[p][n];
    for (int cc = p + 1; cc < n; cc++) s -= A[p][cc] * x[cc];
    x[p] = A[p][p] > 0.0 ? s / A[p][p] : 0.0;
  }
  for (int r = 0; r < n; r++) d[B.off + r] = (S)x[r];
}

// ---------------------------------------------------------------------------
// Jacobi-preconditioned conjugate gradients on a coarsest level too large for
// the direct solve (a hierarchy cut at a few levels).  A is symmetric (each
// face conductance enters both of its cells) and positive definite wherever
// a BC diagonal reaches the block.  Vectors in S, inner products and the
// step scalars in FP64.  Each iteration is two kernels; each writes one
// partial inner product per work-group, and every work-group of the next
// kernel adds all of them in the same fixed order (no atomics, no device
// fences, the same bits on every run and in every group).
// ---------------------------------------------------------------------------
// Sum of v over the 256 threads of the work-group (valid in every thread).
__device__ __forceinline__ double cg_group_sum(double v, double* sm) {
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  __syncthreads();
  if ((threadIdx.x & 63) == 0) sm[threadIdx.x >> 6] = v;
  __syncthreads();
  return (sm[0] + sm[1]) + (sm[2] + sm[3]);
}
// Sum of the n partials of the previous kernel, in a fixed order.
__device__ __forceinline__ double cg_all_sum(const double* __restrict__ part, int n, double* sm) {
  double t = 0.0;
  for (int q = threadIdx.x; q < n; q += 256) t += part[q];
  return cg_group_sum(t, sm);
}

// (A v)_g with v(g') = z(g') + beta p(g'); *vc = v(g)
template <typename S>
__device__ __forceinline__ S cg_Ap(const COp<S>& o, const S* __restrict__ z, const S* __restrict__ p, S beta,
                                   const CBlk& B, int g, int i, int j, int k, S* vc) {
  const int sy = B.n[0], sz = B.n[0] * B.n[1];
  auto v = [&](int q) -> S { return z[q] + beta * p[q]; };
  const S c0 = v(g);
  *vc = c0;
  S a = o.D[g] * c0;
  if (i > 0) a += o.cx[g] * (c0 - v(g - 1));
  if (i < B.n[0] - 1) a += o.cx[g + 1] * (c0 - v(g + 1));
  if (j > 0) a += o.cy[g] * (c0 - v(g - sy));
  if (j < B.n[1] - 1) a += o.cy[g + sy] * (c0 - v(g + sy));
  if (k > 0) a += o.cz[g] * (c0 - v(g - sz));
  if (k < B.n[2] - 1) a += o.cz[g + sz] * (c0 - v(g + sz));
  return a;
}

// r = b - A x, invd = 1/diag(A), z = invd r, p = z; partials of r . z
template <typename S>
__global__ __launch_bounds__(256) void k_cg_start(COp<S> o, const S* __restrict__ x, S* __restrict__ r,
                                                  S* __restrict__ z, S* __restrict__ p, S* __restrict__ invd,
                                                  double* __restrict__ prz) {
  __shared__ double sm[4];
  const int g = blockIdx.x * 256 + threadIdx.x;
  double v = 0.0;
  if (g < o.ncell) {
    const int bi = cmg_find_block(o.blk, o.nb, g);
    const CBlk B = o.blk[bi];
    const CCell c = cmg_cell(B, bi, g);
    S dg;
    const S ax = cmg_Ad(o, x, B, g, c.i, c.j, c.k, &dg);
    const S id = dg > (S)0 ? (S)1 / dg : (S)0;
    const S rr = o.b[g] - ax;
    const S zz = id * rr;
    r[g] = rr;
    z[g] = zz;
    p[g] = zz;
    invd[g] = id;
    v = (double)rr * (double)zz;
  }
  v = cg_group_sum(v, sm);
  if (threadIdx.x == 0) prz[blockIdx.x] = v;
}

// Iteration it: rz = r . z (from prz), beta = rz / rz of the previous
// iteration (rzs[(it+1)%2]; 0 on the first), p' = z + beta p, q = A p';
// partials of p' . q; work-group 0 keeps rz in rzs[it%2].
template <typename S>
__global__ __launch_bounds__(256) void k_cg_ap(COp<S> o, int it, const S* __restrict__ z, const S* __restrict__ p,
                                               S* __restrict__ pn, S* __restrict__ q,
                                               const double* __restrict__ prz, double* __restrict__ ppq,
                                               double* __restrict__ rzs) {
  __shared__ double sm[4];
  const double rz = cg_all_sum(prz, (int)gridDim.x, sm);
  const double rzp = rzs[(it + 1) % 2];
  const S beta = it > 0 && rzp > 0.0 ? (S)(rz / rzp) : (S)0;
  const int g = blockIdx.x * 256 + threadIdx.x;
  double v = 0.0;
  if (g < o.ncell) {
    const int bi = cmg_find_block(o.blk, o.nb, g);
    const CBlk B = o.blk[bi];
    const CCell c = cmg_cell(B, bi, g);
    S pc;
    const S a = cg_Ap(o, z, p, beta, B, g, c.i, c.j, c.k, &pc);
    pn[g] = pc;
    q[g] = a;
    v = (double)pc * (double)a;
  }
  v = cg_group_sum(v, sm);
  if (threadIdx.x == 0) {
    ppq[blockIdx.x] = v;
    if (blockIdx.x == 0) rzs[it % 2] = rz;
  }
}

// alpha = rz / (p' . q); x += alpha p', r -= alpha q, z = invd r; partials
// of the new r . z into przn
template <typename S>
__global__ __launch_bounds__(256) void k_cg_update(int n, S* __restrict__ x, const S* __restrict__ p,
                                                   S* __restrict__ r, const S* __restrict__ q,
                                                   const S* __restrict__ invd, S* __restrict__ z,
                                                   const double* __restrict__ prz, const double* __restrict__ ppq,
                                                   double* __restrict__ przn) {
  __shared__ double sm[4];
  const double rz = cg_all_sum(prz, (int)gridDim.x, sm);
  const double pq = cg_all_sum(ppq, (int)gridDim.x, sm);
  const S alpha = pq > 0.0 ? (S)(rz / pq) : (S)0;
  const int g = blockIdx.x * 256 + threadIdx.x;
  double v = 0.0;
  if (g < n) {
    x[g] = x[g] + alpha * p[g];
    const S rr = r[g] - alpha * q[g];
    const S zz = invd[g] * rr;
    r[g] = rr;
    z[g] = zz;
    v = (double)rr * (double)zz;
  }
  v = cg_group_sum(v, sm);
  if (threadIdx.x == 0) przn[blockIdx.x] = v;
}

template <typename S>
__global__ __launch_bounds__(256) void k_cmg_fill(S* __restrict__ p, int n, S v) {
  const int g = blockIdx.x * 256 + threadIdx.x;
  if (g < n) p[g] = v;
}

// ---------------------------------------------------------------------------
// Finest level: FP64 defect, conductances and BC diagonal from the padded
// arena of the explicit step (ghosts filled, BC values in bcv).
// ---------------------------------------------------------------------------

// kappa(face average T, left material) * normal weight of the face between q
// and its +e_D neighbour.
template <int D>
__device__ __forceinline__ double face_cond(const StageArgs<double>& a, const BlockInfo& B, long q) {
  const long s = D == 0 ? 1 : (D == 1 ? B.sy : B.sz);
  int bad = 0;
  const double kap = tab_interp(a.tab.kap + (long)a.mat[q] * a.tab.nT, 0.5 * (a.T[q] + a.T[q + s]), a.tab.Ti1,
                                a.tab.Ti2, &bad);
  return kap * a.W[3 * D + D][q];
}

// -d(bcv)/dT_cell of one BC entry with kappa (and the radiative K) frozen:
// 302: 2 A kappa (M_dir . N); 303: A / H; 304/305: A / H at the converged
// wall temperature; other types: 0.
template <int NDIR>
__global__ void k_bcderiv(StageArgs<double> a, double* __restrict__ bcd, int n) {
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= n) return;
  const BCRec& r = a.bc[e];
  double d = 0.0;
  if (r.type == 302 || r.type == 303 || r.type == 304 || r.type == 305) {
    const long p = r.cell;
    const double Tc = a.T[p];
    int bad = 0;
    const double kappa = tab_interp(a.tab.kap + (long)a.mat[p] * a.tab.nT, Tc, a.tab.Ti1, a.tab.Ti2, &bad);
    const double* M = r.M;
    const double mn = M[r.dir - 1] * r.N[0] + M[r.dir - 1 + 3] * r.N[1] + M[r.dir - 1 + 6] * r.N[2];
    if (r.type == 302) {
      d = 2.0 * r.A * kappa * mn;
    } else if (r.type == 303) {
      const double H = 1.0 / (2.0 * kappa * mn) + 1.0 / r.hconv;
      d = r.A / H;
    } else {
      const double Tref = r.Tref, Tr2 = Tref * Tref, Tr4 = Tr2 * Tr2;
      double Tw = Tc, qw = 0.0, err = 1.0e6, H = 1.0 / (2.0 * kappa * mn);
      int it = 0;
      while (err > 1.0e-3 && it++ < 1000) {
        const double Tw2 = Tw * Tw;
        const double K = Tw != Tref ? SIGMA_SB * r.eps * (Tw2 * Tw2 - Tr4) / (Tw - Tref)
                                    : 4.0 * SIGMA_SB * r.eps * Tr2 * Tref;
        H = r.type == 304 ? 1.0 / (2.0 * kappa * mn) + 1.0 / K : 1.0 / (2.0 * kappa * mn) + 1.0 / (K + r.hconv);
        const double q_wall = K * (Tref - Tw);
        err = fabs(qw - q_wall);
        qw = q_wall;
        Tw = Tref - 1.0 / (H * K) * (Tref - Tc);
        if (err != err) break;
      }
      d = r.A / H;
    }
  }
  bcd[e] = d;
}

template <int NDIR, bool TANG, typename S>
__global__ __launch_bounds__(TILE_T) void k_defect(StageArgs<double> a, const double* __restrict__ bcd,
                                                   const CBlk* __restrict__ cblk, S* __restrict__ b, S* __restrict__ cx,
                                                   S* __restrict__ cy, S* __restrict__ cz, S* __restrict__ D,
                                                   double* __restrict__ partials) {
  const int wg = xcd_remap(blockIdx.x, gridDim.x);
  const Tile t = a.tiles[wg];
  const BlockInfo B = a.blk[t.b];
  const int i = t.i0 + (threadIdx.x % TILE_X), j = t.j0 + (threadIdx.x / TILE_X), k = t.k0;
  double bsq = 0.0;
  if (i <= B.ni && j <= B.nj) {
    const long p = B.base + (i - 1) + (long)(j - 1) * B.sy + (long)(k - 1) * B.sz;
    double Rv = 0;
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
    Rv = apply_bc_faces(a, B, i, j, k, Rv);
    const double bv = -Rv + a.qvol[p] * a.vol[p];
    const CBlk& C = cblk[t.b];
    const int g = C.off + (i - 1) + C.n[0] * ((j - 1) + C.n[1] * (k - 1));
    b[g] = (S)bv;
    cx[g] = (S)(i > 1 ? face_cond<0>(a, B, p - 1) : 0.0);
    cy[g] = (S)(j > 1 ? face_cond<1>(a, B, p - B.sy) : 0.0);
    cz[g] = (S)((NDIR == 3 && k > 1) ? face_cond<2>(a, B, p - B.sz) : 0.0);
    double dd = 0.0;
    const int fa = B.face_active;
    if (fa) {
      if ((fa & 1) && i == 1) dd += bcd[a.fmap[B.fmap_off[0] + (j - 1) + B.nj * (k - 1)]];
      if ((fa & 2) && i == B.ni) dd += bcd[a.fmap[B.fmap_off[1] + (j - 1) + B.nj * (k - 1)]];
      if ((fa & 4) && j == 1) dd += bcd[a.fmap[B.fmap_off[2] + (i - 1) + B.ni * (k - 1)]];
      if ((fa & 8) && j == B.nj) dd += bcd[a.fmap[B.fmap_off[3] + (i - 1) + B.ni * (k - 1)]];
      if ((fa & 16) && k == 1) dd += bcd[a.fmap[B.fmap_off[4] + (i - 1) + B.ni * (j - 1)]];
      if ((fa & 32) && k == B.nk) dd += bcd[a.fmap[B.fmap_off[5] + (i - 1) + B.ni * (j - 1)]];
    }
    D[g] = (S)dd;
    bsq = bv * bv;
  }
  block_sum(bsq, partials, wg);
}

// T += d on the interior cells of the finest level.
template <typename S>
__global__ __launch_bounds__(TILE_T) void k_cmg_apply(const BlockInfo* __restrict__ blk, const Tile* __restrict__ tiles,
                                                      const CBlk* __restrict__ cblk, const S* __restrict__ d,
                                                      double* __restrict__ T) {
  const int wg = xcd_remap(blockIdx.x, gridDim.x);
  const Tile t = tiles[wg];
  const BlockInfo B = blk[t.b];
  const int i = t.i0 + (threadIdx.x % TILE_X), j = t.j0 + (threadIdx.x / TILE_X), k = t.k0;
  if (i > B.ni || j > B.nj) return;
  const long p = B.base + (i - 1) + (long)(j - 1) * B.sy + (long)(k - 1) * B.sz;
  const CBlk& C = cblk[t.b];
  const int g = C.off + (i - 1) + C.n[0] * ((j - 1) + C.n[1] * (k - 1));
  T[p] = T[p] + (double)d[g];
}

}  // namespace fuss
