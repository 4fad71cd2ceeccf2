// fuss_gpu.hip -- MI355X (gfx950) explicit solver step of FUSS behind the
// C-ABI of include/fuss_gpu.h.
//
// One HIP stream per context.  Per multigrid level all blocks share one
// arena per field; kernels walk a work-list of 64x4 cell tiles covering every
// block, so uneven multi-block levels run in one launch per pass.
//
// Passes of one Explicit_Step (src/lib/numerics/time/explicit/Mod_Explicit.f90:10-111):
//   k_dt        local dt from T^n + global min                (Mod_dt.f90:10-84)
//   per stage:
//     k_ghost   layer-1 ghosts: mirror or block-to-block copy (Lib_Ghost.f90:11-189)
//     k_stage   gather-form interior fluxes + BC fluxes + RK   (Mod_Fluxes.f90, Mod_BC_Fluxes.f90,
//               update (or the scaled residual when IRS is on)  Lib_Newstate.f90)
//     k_irs     ndir line passes of 2 Jacobi sweeps, the last   (Lib_IRS.f90)
//               fused with the RK update
//   k_reduce    Sum (T^{n+1}-T^n)^2, deterministic two-pass     (Mod_Diagnostic.f90:13-49)
// Copy_State is free: T^n, and the two stage buffers rotate.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/fuss_gpu.h"
#include <mutex>
#include <type_traits>

#include "fuss_host.h"

// ------------------------------------------------------------------------
// kernels: geometry (setup)
// ------------------------------------------------------------------------
struct GeomArgs {
  const BlockInfo* blk;
  const Tile* tiles;
  const double* nodes;   // node arena, 3 per node
  double* M;             // 9 per padded cell
  double* vol;
  double* invL2;         // sum_{d<=ndir} 1/dl_d^2 (Mod_dt.f90:280-283)
  int ndir;
  int* status;
};

__device__ __forceinline__ const double* node_at(const double* nodes, const BlockInfo& B, int i, int j,
                                                 int k) {
  return nodes + 3 * (B.node_off + i + (long)(B.ni + 1) * (j + (long)(B.nj + 1) * k));
}

// Setup_Metrics interior loop, src/lib/numerics/space/Mod_Metrics.f90:23-41
__global__ __launch_bounds__(TILE_T) void k_cell_metrics(GeomArgs a) {
  const Tile t = a.tiles[blockIdx.x];
  const BlockInfo B = a.blk[t.b];
  const int i = t.i0 + (threadIdx.x % TILE_X), j = t.j0 + (threadIdx.x / TILE_X), k = t.k0;
  if (i > B.ni || j > B.nj) return;
  const long p = B.base + (i - 1) + (long)(j - 1) * B.sy + (long)(k - 1) * B.sz;
  double m[9], dl[3], v;
  const int e = metric_tensor(node_at(a.nodes, B, i - 1, j - 1, k - 1), node_at(a.nodes, B, i - 1, j - 1, k),
                              node_at(a.nodes, B, i - 1, j, k - 1), node_at(a.nodes, B, i - 1, j, k),
                              node_at(a.nodes, B, i, j - 1, k - 1), node_at(a.nodes, B, i, j - 1, k),
                              node_at(a.nodes, B, i, j, k - 1), node_at(a.nodes, B, i, j, k), m, dl, &v);
  if (e) flag_error(a.status, FUSS_EMETRIC, t.b, i, j, k);
  for (int q = 0; q < 9; q++) a.M[9 * p + q] = m[q];
  a.vol[p] = v;
  double inv = 0.0;
  {
#pragma clang fp contract(off)
    for (int d = 0; d < a.ndir; d++) inv = inv + 1.0 / (dl[d] * dl[d]);
  }
  a.invL2[p] = inv;
}

struct FaceArgs {
  BlockInfo B;
  int bidx;
  int D;                 // direction 0..2
  double sign;
  const double* nodes;
  const double* M;
  double* F;             // 4 per padded cell: N0,N1,N2,A of the face (cell, cell+1 in D)
  double* W[3];          // weights of direction D
  int* tang;             // set when a tangential weight is nonzero
};

// Compute_Norm_Area, Lib_Metrics.f90:69-247, plus the flux weights of the
// interior faces (Lib_Diffusive.f90:34-40 regrouped per computational direction).
__global__ void k_face_metrics(FaceArgs a) {
  const BlockInfo& B = a.B;
  const int D = a.D;
  const long ext0 = B.ni + (D == 0), ext1 = B.nj + (D == 1), ext2 = B.nk + (D == 2);
  const long idx = blockIdx.x * (long)blockDim.x + threadIdx.x;
  if (idx >= ext0 * ext1 * ext2) return;
  const int ii = (int)(idx % ext0), jj = (int)((idx / ext0) % ext1), kk = (int)(idx / (ext0 * ext1));
  // Fortran face indices: dir1 i=0..ni, j=1..nj, k=1..nk, etc.
  const int i = D == 0 ? ii : ii + 1, j = D == 1 ? jj : jj + 1, k = D == 2 ? kk : kk + 1;
  const double *p, *q, *r, *s;
  if (D == 0) {
    p = node_at(a.nodes, B, i, j, k); q = node_at(a.nodes, B, i, j - 1, k - 1);
    r = node_at(a.nodes, B, i, j - 1, k); s = node_at(a.nodes, B, i, j, k - 1);
  } else if (D == 1) {
    p = node_at(a.nodes, B, i - 1, j, k); q = node_at(a.nodes, B, i, j, k - 1);
    r = node_at(a.nodes, B, i, j, k); s = node_at(a.nodes, B, i - 1, j, k - 1);
  } else {
    p = node_at(a.nodes, B, i - 1, j, k); q = node_at(a.nodes, B, i, j - 1, k);
    r = node_at(a.nodes, B, i - 1, j - 1, k); s = node_at(a.nodes, B, i, j, k);
  }
  double N[3], A;
  face_normal(p, q, r, s, a.sign, N, &A);
  // the face (i,j,k) of direction D lies between cells c and c+e_D, c = (i,j,k)
  // with the D index as given (0 = ghost side of the min face)
  const long c = B.base + (i - 1) + (long)(j - 1) * B.sy + (long)(k - 1) * B.sz;
  a.F[4 * c + 0] = N[0];
  a.F[4 * c + 1] = N[1];
  a.F[4 * c + 2] = N[2];
  a.F[4 * c + 3] = A;
  const int idD = D == 0 ? i : (D == 1 ? j : k);
  const int nD = D == 0 ? B.ni : (D == 1 ? B.nj : B.nk);
  if (idD >= 1 && idD < nD) {
#pragma clang fp contract(off)
    const long st = D == 0 ? 1 : (D == 1 ? (long)B.sy : B.sz);
    const double* M1 = a.M + 9 * c;
    const double* M2 = a.M + 9 * (c + st);
    double Mb[9];
    for (int e = 0; e < 9; e++) Mb[e] = 0.5 * (M1[e] + M2[e]);
    int tg = 0;
    for (int cc = 0; cc < 3; cc++) {
      const double w = A * (Mb[cc] * N[0] + Mb[cc + 3] * N[1] + Mb[cc + 6] * N[2]);
      a.W[cc][c] = w;
      if (cc != D && w != 0.0) tg = 1;
    }
    if (tg) atomicOr(a.tang, 1);
  }
}

// Deviation of the per-axis factors (RectAxes) of the rectilinear blocks
// from the stored metrics, with the products the stage kernel forms: *dev =
// the largest relative deviation of a volume or interior normal weight (bits
// of a non-negative double, so the integer maximum is the numeric one),
// *nbad += cells whose sum 1/dl^2 differs in any bit.
__global__ __launch_bounds__(TILE_T) void k_rect_check(const BlockInfo* __restrict__ blk, const Tile* __restrict__ tiles,
                                                       const double* __restrict__ rect, const double* __restrict__ vol,
                                                       const double* __restrict__ invL2, const double* __restrict__ W0,
                                                       const double* __restrict__ W4, const double* __restrict__ W8,
                                                       unsigned long long* dev, int* nbad) {
  const Tile t = tiles[blockIdx.x];
  const BlockInfo B = blk[t.b];
  if (B.roff < 0) return;
  const int i = t.i0 + threadIdx.x % TILE_X, j = t.j0 + threadIdx.x / TILE_X, k = t.k0;
  double d = 0.0;
  int bad = 0;
  if (i <= B.ni && j <= B.nj) {
    const RectAxes r = rect_axes(rect, B);
    const long p = B.base + (i - 1) + (long)(j - 1) * B.sy + (long)(k - 1) * B.sz;
    auto rel = [](double ref, double v) { return v == ref ? 0.0 : fabs(v - ref) / fabs(ref); };
    const double axy = r.ax[i] * r.ay[j];
    d = rel(vol[p], axy * r.az[k]);
    if (i < B.ni) d = fmax(d, rel(W0[p], (r.fx[i] * r.ay[j]) * r.az[k]));
    if (j < B.nj) d = fmax(d, rel(W4[p], (r.fy[j] * r.ax[i]) * r.az[k]));
    if (k < B.nk) d = fmax(d, rel(W8[p], r.fz[k] * axy));
    if (!(d <= 1.0)) d = 1.0;  // NaN or a gross mismatch
    bad = invL2[p] != (r.qx[i] + r.qy[j]) + r.qz[k];
  }
  for (int o = 32; o > 0; o >>= 1) {
    d = fmax(d, __shfl_xor(d, o, 64));
    bad += __shfl_xor(bad, o, 64);
  }
  if ((threadIdx.x & 63) == 0) {
    atomicMax(dev, (unsigned long long)__double_as_longlong(d));
    if (bad) atomicAdd(nbad, bad);
  }
}

// Fill geometry of BC records from the cell metrics and faces.
__global__ void k_bc_geometry(BCRec* bc, int n, const BlockInfo* blk, const double* M, const double* F0,
                              const double* F1, const double* F2) {
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= n) return;
  BCRec& r = bc[e];
  const BlockInfo& B = blk[r.b];
  // Face_Index, Lib_BC_Fluxes.f90:27-45: the face of a min face lies at cell-1
  int fi = r.i, fj = r.j, fk = r.k;
  if (r.f == 1) fi -= 1;
  if (r.f == 3) fj -= 1;
  if (r.f == 5) fk -= 1;
  const long fc = B.base + (fi - 1) + (long)(fj - 1) * B.sy + (long)(fk - 1) * B.sz;
  const double* F = r.dir == 1 ? F0 : (r.dir == 2 ? F1 : F2);
  r.N[0] = F[4 * fc + 0];
  r.N[1] = F[4 * fc + 1];
  r.N[2] = F[4 * fc + 2];
  r.A = F[4 * fc + 3];
  for (int q = 0; q < 9; q++) r.M[q] = M[9 * r.cell + q];
  {
#pragma clang fp contract(off)
    for (int rr = 0; rr < 3; rr++)
      r.mc[rr] = r.Mg[rr] * r.N[0] + r.Mg[rr + 3] * r.N[1] + r.Mg[rr + 6] * r.N[2];
  }
}

// Setup_BC per entry on the device (the set-up of levels without chimera
// entries): the entry's record (cell, layer-1 ghost and its source, donor
// cells of Tg(3:6) and of the IRS ghost, the BC values, the connection's
// ghost metric Mg(1) from the device nodes), its (ghost, source) pair and its
// face-map slot, as build_bc forms them on the host.  Checks of build_bc
// become error keys (entry << 8 | code, the smallest one wins); the face
// activity and connection bits go to bbits[2 b], bbits[2 b + 1], the level's
// x / y / z connection and ghost-read flags to lflags[0..3].
struct BcBuildArgs {
  const int* ib;           // 15 integers per entry
  const double* rb;        // 6 values per entry
  const int* ser;          // time series per entry, or null
  int n, nb;
  const BlockInfo* blk;
  const double* nodes;
  BCRec* rec;
  int2* gm;
  int* fmap;               // -1 before the build
  int* bbits;
  int* lflags;
  unsigned long long* err;
};

__device__ __forceinline__ long cidx_d(const BlockInfo& B, int i, int j, int k) {
  return B.base + (i - 1) + (long)(j - 1) * B.sy + (long)(k - 1) * B.sz;
}

__global__ void k_build_bc(BcBuildArgs a) {
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= a.n) return;
  auto fail = [&](int code) { atomicMin(a.err, ((unsigned long long)e << 8) | (unsigned)code); };
  const int* q = a.ib + 15 * (long)e;
  const double* r = a.rb + 6 * (long)e;
  // grouping: (block, face) non-decreasing (Mod_BC_Fluxes.f90:26-33)
  if (e > 0) {
    const int* p = q - 15;
    if (q[0] < p[0] || (q[0] == p[0] && q[4] < p[4])) atomicMin(a.err, (unsigned long long)FUSS_EBC);
  }
  const int b = q[0] - 1, i = q[1], j = q[2], k = q[3], f = q[4], type = q[5];
  if (b < 0 || b >= a.nb || f < 1 || f > 6) return fail(FUSS_EBC);
  const BlockInfo B = a.blk[b];
  if (i < 1 || i > B.ni || j < 1 || j > B.nj || k < 1 || k > B.nk) return fail(FUSS_EBC);
  const bool on = (f == 1 && i == 1) || (f == 2 && i == B.ni) || (f == 3 && j == 1) || (f == 4 && j == B.nj) ||
                  (f == 5 && k == 1) || (f == 6 && k == B.nk);
  if (!on) return fail(FUSS_EBC);
  const int mi = f <= 2 ? (j - 1) + B.nj * (k - 1) : (f <= 4 ? (i - 1) + B.ni * (k - 1) : (i - 1) + B.ni * (j - 1));
  const int fpos = B.fmap_off[f - 1] + mi;
  if (atomicCAS(a.fmap + fpos, -1, e) != -1) return fail(FUSS_EBC);
  if (type == 102) return fail(FUSS_EUNSUPPORTED);  // chimera levels are built on the host
  BCRec R;
  memset(&R, 0, sizeof(R));
  R.type = type; R.f = f; R.dir = (f + 1) / 2; R.series = a.ser ? a.ser[e] : -1;
  R.b = b; R.i = i; R.j = j; R.k = k;
  R.cell = cidx_d(B, i, j, k);
  R.fpos = fpos;
  R.qw = r[0]; R.Tw = r[1]; R.hconv = r[2]; R.Tref = r[4]; R.eps = r[5];
  R.gdst = cidx_d(B, i - guide(f, 0), j - guide(f, 1), k - guide(f, 2));
  R.gsrc = R.cell;
  R.rdonor = -1;
  for (int t = 0; t < 4; t++) R.dn[t] = R.cell;
  if (type == 101) {
    const int bs = q[6] - 1, Is = q[7], Js = q[8], Ks = q[9], Fs = q[10];
    if (bs < 0 || bs >= a.nb || Fs < 1 || Fs > 6) return fail(FUSS_EBC);
    const BlockInfo S = a.blk[bs];
    if (Is < 1 || Is > S.ni || Js < 1 || Js > S.nj || Ks < 1 || Ks > S.nk) return fail(FUSS_EBC);
    R.gsrc = cidx_d(S, Is, Js, Ks);
    R.rdonor = R.gsrc;
    const int d11 = q[11], d12 = q[12], d21 = q[13], d22 = q[14];
    for (int t = 0; t < 4; t++) {
      // Tg(3:6): the donor's neighbours along its face's two tangents
      const int sg = (t & 1) ? 1 : -1, da = t < 2 ? d11 : d21, db = t < 2 ? d12 : d22;
      int pi = Is, pj = Js, pk = Ks;
      if (Fs <= 2) { pj += sg * da; pk += sg * db; }
      else if (Fs <= 4) { pi += sg * da; pk += sg * db; }
      else { pi += sg * da; pj += sg * db; }
      if (pi < 0 || pi > S.ni + 1 || pj < 0 || pj > S.nj + 1 || pk < 0 || pk > S.nk + 1) return fail(FUSS_EBC);
      R.dn[t] = cidx_d(S, pi, pj, pk);
    }
    const int me = connect_metric_nodes(a.nodes + 3 * B.node_off, B.ni, B.nj, B.nk, a.nodes + 3 * S.node_off, S.ni,
                                        S.nj, S.nk, q, R.Mg);
    if (me) return fail(me == 1 ? FUSS_EBC : FUSS_EMETRIC);
  }
  a.rec[e] = R;
  a.gm[e] = make_int2((int)R.gdst, (int)R.gsrc);
  const bool active = type == 101 || (type >= 300 && type <= 305);
  if (active) atomicOr(a.bbits + 2 * b, 1 << (f - 1));
  if (type == 101) {
    atomicOr(a.bbits + 2 * b + 1, 1 << (f - 1));
    atomicOr(a.lflags + (f - 1) / 2, 1);
  }
  if (type == 101 || type == 300 || type == 302) atomicOr(a.lflags + 3, 1);
}

// every face cell of the level has its entry
__global__ void k_fmap_check(const int* fmap, long n, int nent, unsigned long long* err) {
  for (long q = blockIdx.x * (long)blockDim.x + threadIdx.x; q < n; q += (long)gridDim.x * blockDim.x)
    if (fmap[q] < 0) atomicMin(err, ((unsigned long long)nent << 8) | (unsigned)FUSS_EBC);
}

// Level-1 fields of block B from the caller's arrays (i fastest, nc cells):
// T, qvol and the 1-based material index into the padded arena; the material
// range is checked (err = 1) and its min / max kept in mm[0], mm[1].
__global__ void k_scatter_fields(BlockInfo B, const double* T, const double* qv, const double* mid, double* Td,
                                 double* qd, int* md, int nmat, int* mm, int* err) {
  const long nc = (long)B.ni * B.nj * B.nk;
  int lo = 0x7fffffff, hi = -0x7fffffff;
  for (long s = blockIdx.x * (long)blockDim.x + threadIdx.x; s < nc; s += (long)gridDim.x * blockDim.x) {
    const int i = (int)(s % B.ni), j = (int)((s / B.ni) % B.nj), k = (int)(s / ((long)B.ni * B.nj));
    const long p = B.base + i + (long)j * B.sy + (long)k * B.sz;
    Td[p] = T[s];
    qd[p] = qv[s];
    const int m = (int)mid[s] - 1;
    if (m < 0 || m >= nmat) *err = 1;
    md[p] = m;
    lo = min(lo, m);
    hi = max(hi, m);
  }
  atomicMin(mm, lo);
  atomicMax(mm + 1, hi);
}

// matIDg of the connection entries: the material of the donor cell
// (Ghost_Connection copies it, Fill_matIDg, Lib_Ghost.f90:149-153, :436-470)
__global__ void k_set_matg(BCRec* bc, int n, const int* mat) {
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e < n && bc[e].type == 101) bc[e].matg = mat[bc[e].gsrc];
}

// blocks whose qvol differs from the value of their cell (1,1,1): nonu[b] = 1
// (work-group (b, y) scans planes y, y + gridDim.y, ... of block b)
__global__ void k_qvol_nonuniform(const BlockInfo* blk, const double* q, int* nonu) {
  const BlockInfo B = blk[blockIdx.x];
  const double q0 = q[B.base];
  const int nij = B.ni * B.nj;
  int bad = 0;
  for (int k = blockIdx.y; k < B.nk && !bad; k += gridDim.y)
    for (int e = threadIdx.x; e < nij; e += blockDim.x)
      if (q[B.base + e % B.ni + (long)(e / B.ni) * B.sy + (long)k * B.sz] != q0) bad = 1;
  if (bad) nonu[blockIdx.x] = 1;
}

__global__ void k_fill_i32(int* p, long n, int v) {
  for (long q = blockIdx.x * (long)blockDim.x + threadIdx.x; q < n; q += (long)gridDim.x * blockDim.x) p[q] = v;
}

// ------------------------------------------------------------------------
// kernels: ghost fill (per stage)
// ------------------------------------------------------------------------
// Ghost_Connection / Ghost_Symmetry (Lib_Ghost.f90:132-189), first layer:
// the explicit step's stencils read no cell beyond it.
// g[e] = (ghost, source) arena indices of entry e, ghost < 0: none.
template <typename R>
__global__ void k_ghost(R* __restrict__ T, const int2* __restrict__ g, int n, const int* __restrict__ halt) {
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= n || (halt && *halt)) return;
  const int2 q = g[e];
  if (q.x >= 0) stu(T, (unsigned)q.x, ldu((const R*)T, (unsigned)q.y));
}

// Ghost_Chimera (Lib_Ghost.f90:192-245), first layer: the ghost's enthalpy is
// the volume-fraction-weighted sum of its donors' (co_H at each donor's
// material and T, summed in donor order), its T the inversion (co_T) with the
// last donor's material.  r = (ghost, first donor, donors, material); the
// second layer is read by no term of the step and is not formed.
__global__ void k_chimera(double* __restrict__ T, const int4* __restrict__ r, const int* __restrict__ don,
                          const double* __restrict__ vf, int n, const int* __restrict__ mat, Tab<double> tab,
                          int* status, const int* __restrict__ halt) {
#pragma clang fp contract(off)
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= n || (halt && *halt)) return;
  const int4 q = r[e];
  double consg = 0.0;
  int bad = 0;
  for (int t = q.y; t < q.y + q.z; t++) {
    const int p = don[t];
    const double consi = tab_interp(tab.h + (long)mat[p] * tab.nT, T[p], tab.Ti1, tab.Ti2, &bad);
    consg = consg + consi * vf[t];
  }
  T[q.x] = tab_inverse_search(tab.h + (long)q.w * tab.nT, consg, tab.Ti1, tab.Ti2, &bad);
  if (bad) flag_error(status, FUSS_ETABLE, -1, 0, 0, 0);
}

// Precision conversion of a padded field (FP32 explicit steps).
__global__ __launch_bounds__(256) void k_d2f(const double* __restrict__ src, float* __restrict__ dst, long n) {
  for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x)
    dst[i] = (float)src[i];
}
__global__ __launch_bounds__(256) void k_f2d(const float* __restrict__ src, double* __restrict__ dst, long n) {
  for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x)
    dst[i] = (double)src[i];
}

// Compute_dt, Mod_dt.f90:26-84
template <typename R>
__global__ __launch_bounds__(TILE_T) void k_dt(DtArgs<R> a) {
  if (a.cnt[1]) return;
  const int wg = xcd_remap(blockIdx.x, gridDim.x);
  dt_tile(a, wg, a.cnt[0] + 1);
}

template <int NDIR, typename R>
__global__ void k_bcflux(StageArgs<R> a, const BCHot* __restrict__ hot, int n) {
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= n || (a.halt && *a.halt)) return;
  const int type = hot[e].type;
  if (!(type == 101 || type == 102 || (type >= 300 && type <= 305)) || (type == 101 && a.conn_split)) return;
  const R v = bc_entry_value<NDIR>(a, hot, e, *a.clk);
  a.bcv[e] = v;
  a.bcf[hot[e].fpos] = v;
}

// Connection entries (101) of the compact records: doubles M(9), N(3),
// Mg(1).N (3), A; ints cell, ghost donor, Tg(3:6) donors (4), face | matIDg
// << 3, sy, sz, entry, face-map position, block, i | j << 16, k.
constexpr int CONN_ND = 16, CONN_NI = 15;

// BC_Connection of every 101 entry (the surface pass's connection part, one
// thread per entry, records read field by field: coalesced): the same
// arithmetic as bc_value's, its value to bcv and bcf.
template <int NDIR, typename R>
__global__ void k_connflux(StageArgs<R> a, const double* __restrict__ cd, const int* __restrict__ ci, int n) {
  const int q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= n || (a.halt && *a.halt)) return;
  double M[9], Nr[3], mc[3];
  for (int f = 0; f < 9; f++) M[f] = cd[(long)f * n + q];
  for (int f = 0; f < 3; f++) Nr[f] = cd[(long)(9 + f) * n + q];
  for (int f = 0; f < 3; f++) mc[f] = cd[(long)(12 + f) * n + q];
  const double A = cd[15L * n + q];
  const long p = ci[q], gsrc = ci[n + q];
  const long dn[4] = {ci[2L * n + q], ci[3L * n + q], ci[4L * n + q], ci[5L * n + q]};
  const int fm = ci[6L * n + q];
  const long sy = ci[7L * n + q], sz = ci[8L * n + q];
  const int e = ci[9L * n + q], fpos = ci[10L * n + q];
  int bad = 0;
  const R v = conn_value<NDIR, R>(a.T, p, sy, sz, fm & 7, M, Nr, mc, A, gsrc, dn,
                                  a.tab.kap + (long)a.mat[p] * a.tab.nT, a.tab.kap + (long)(fm >> 3) * a.tab.nT,
                                  a.tab.Ti1, a.tab.Ti2, &bad);
  if (bad) {
    const int ij = ci[12L * n + q];
    flag_error(a.status, FUSS_ETABLE, ci[11L * n + q], ij & 0xffff, ij >> 16, ci[13L * n + q]);
  }
  a.bcv[e] = v;
  a.bcf[fpos] = v;
}

// the compact records of the connection entries listed in ent
__global__ void k_conn_build(const BCRec* __restrict__ bc, const int* __restrict__ ent, int n,
                             const BlockInfo* __restrict__ blk, double* cd, int* ci) {
  const int q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= n) return;
  const int e = ent[q];
  const BCRec& r = bc[e];
  const BlockInfo& B = blk[r.b];
  for (int f = 0; f < 9; f++) cd[(long)f * n + q] = r.M[f];
  for (int f = 0; f < 3; f++) cd[(long)(9 + f) * n + q] = r.N[f];
  for (int f = 0; f < 3; f++) cd[(long)(12 + f) * n + q] = r.mc[f];
  cd[15L * n + q] = r.A;
  const int iv[CONN_NI] = {(int)r.cell, (int)r.gsrc, (int)r.dn[0], (int)r.dn[1], (int)r.dn[2], (int)r.dn[3],
                           r.f | (r.matg << 3), B.sy, (int)B.sz, e, r.fpos, r.b, r.i | (r.j << 16), r.k, 0};
  for (int f = 0; f < CONN_NI; f++) ci[(long)f * n + q] = iv[f];
}

// Extract_Properties_From_BC (src/lib/io/IO_Wall.f90:131-239): wall
// temperature and heat flux (Compute_Wall_Properties, Lib_BC_Fluxes.f90:48-58)
// of every entry of types 301-305 at the state a.T (0 for other types), with
// the wall routines' arithmetic: 301 T_wall = T - (G M)(dir)
// (Lib_BC_Fluxes_Wall_Heat.f90:37-52), 302 q_wall = kappa (G M).n
// (Lib_BC_Fluxes_Wall_Temperature.f90:37-66), 303 the conductance balance
// (Lib_BC_Fluxes_Wall_HeatTransfer.f90:44-46, its unset qw as 0), 304/305 the
// last iterate of the radiative fixed point.
template <int NDIR>
__global__ void k_wall(StageArgs<double> a, int n, double* __restrict__ Tw_out, double* __restrict__ qw_out) {
#pragma clang fp contract(off)
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= n) return;
  const BCRec& r = a.bc[e];
  double Tw = 0.0, qw = 0.0;
  if (r.type >= 301 && r.type <= 305) {
    const BlockInfo B = a.blk[r.b];
    const double* __restrict__ T = a.T;
    const long p = r.cell, sy = B.sy, sz = B.sz;
    const int Fm = r.f, dir = r.dir;
    const int modfm2 = 1 - 2 * (Fm % 2), modfm3 = 2 * (2 * (Fm % 2) - 1);
    const double Tc = T[p];
    const double* M = r.M;
    const double* Nr = r.N;
    int bad = 0;
    const double kappa = tab_interp(a.tab.kap + (long)a.mat[p] * a.tab.nT, Tc, a.tab.Ti1, a.tab.Ti2, &bad);
    const double mn = M[dir - 1] * Nr[0] + M[dir - 1 + 3] * Nr[1] + M[dir - 1 + 6] * Nr[2];
    double sv = 0.0;
    if (r.series >= 0)
      sv = series_value(a.ser.per[r.series], a.ser.n[r.series], a.ser.t + a.ser.off[r.series],
                        a.ser.v + a.ser.off[r.series], *a.clk);
    if (r.type == 301) {
      const double q = r.series >= 0 ? sv : r.qw;
      double G[3] = {0.0, 0.0, 0.0};
      G[dir - 1] = -q / (2.0 * kappa * mn) * modfm2;
      if (NDIR == 2) G[2] = 0.0;
      const double grd = G[0] * M[3 * (dir - 1)] + G[1] * M[3 * (dir - 1) + 1] + G[2] * M[3 * (dir - 1) + 2];
      Tw = Tc - grd;
      qw = q;
    } else if (r.type == 302) {
      const double Tv = r.series >= 0 ? sv : r.Tw;
      double G[3] = {0.0, 0.0, 0.0};
      G[dir - 1] = (Tc - Tv) * modfm3;
      if (Fm <= 2) {
        G[1] = (T[p - sy] - T[p + sy]) / 2.0;
        if (NDIR == 3) G[2] = (T[p - sz] - T[p + sz]) / 2.0;
      } else if (Fm <= 4) {
        G[0] = (T[p - 1] - T[p + 1]) / 2.0;
        if (NDIR == 3) G[2] = (T[p - sz] - T[p + sz]) / 2.0;
      } else {
        G[0] = (T[p - 1] - T[p + 1]) / 2.0;
        G[1] = (T[p - sy] - T[p + sy]) / 2.0;
      }
      if (NDIR == 2) G[2] = 0.0;
      double gr[3];
      for (int x = 0; x < 3; x++) gr[x] = G[0] * M[3 * x] + G[1] * M[3 * x + 1] + G[2] * M[3 * x + 2];
      Tw = Tv;
      qw = kappa * (gr[0] * Nr[0] + gr[1] * Nr[1] + gr[2] * Nr[2]);
    } else if (r.type == 303) {
      const double q0 = 0.0;
      Tw = (r.hconv * r.Tref + Tc * 2.0 * kappa * mn + q0) / (r.hconv + 2.0 * kappa * mn);
      qw = r.hconv * (r.Tref - Tw) + q0;
    } else {
      double Tr2 = r.Tref * r.Tref, Tr4 = Tr2 * Tr2, w = Tc, qo = 0.0, err = 1.0e6, q = 0.0;
      int it = 0;
      while (err > 1.0e-3) {
        const double w2 = w * w;
        const double K = SIGMA_SB * r.eps * (w2 * w2 - Tr4) / (w - r.Tref);
        const double H = r.type == 304 ? 1.0 / (2.0 * kappa * mn) + 1.0 / K
                                       : 1.0 / (2.0 * kappa * mn) + 1.0 / (K + r.hconv);
        q = K * (r.Tref - w);
        err = fabs(qo - q);
        qo = q;
        w = r.Tref - 1.0 / (H * K) * (r.Tref - Tc);
        if (++it > 1000 || err != err) {
          flag_error(a.status, FUSS_ESTATE, r.b, r.i, r.j, r.k);
          break;
        }
      }
      Tw = w;
      qw = q;
    }
    if (bad) flag_error(a.status, FUSS_ETABLE, r.b, r.i, r.j, r.k);
  }
  Tw_out[e] = Tw;
  qw_out[e] = qw;
}

// Compact records of the wall entries (after the geometry of the full
// records and the cell materials are on the device).
__global__ void k_bc_hot(const BCRec* __restrict__ bc, int n, const int* __restrict__ mat, BCHot* __restrict__ hot) {
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= n) return;
  const BCRec& r = bc[e];
  BCHot h;
  h.type = r.type; h.f = r.f; h.m = mat[r.cell]; h.series = r.series;
  h.cell = (unsigned)r.cell; h.b = r.b;
  h.fpos = r.fpos; h.pad_ = 0;
  h.A = r.A;
  {
#pragma clang fp contract(off)
    // the dot product bc_value forms for 303/304/305
    h.mn = r.M[r.dir - 1] * r.N[0] + r.M[r.dir - 1 + 3] * r.N[1] + r.M[r.dir - 1 + 6] * r.N[2];
  }
  h.hconv = r.hconv; h.Tref = r.Tref; h.qw = r.qw; h.eps = r.eps;
  hot[e] = h;
}


// ------------------------------------------------------------------------
// kernels: reductions (deterministic: fixed partition and order)
// ------------------------------------------------------------------------
constexpr int RED_CHUNK = 2048;

// Pass 1: work-group b reduces in[b*RED_CHUNK, ...) to out[b].
template <bool MIN>
__global__ __launch_bounds__(256) void k_reduce(const double* __restrict__ in, int n, double* __restrict__ out,
                                                const int* __restrict__ halt = nullptr) {
  if (halt && *halt) return;
  __shared__ double sm[256];
  const int lo = blockIdx.x * RED_CHUNK, hi = min(n, lo + RED_CHUNK);
  double v = MIN ? 1e300 : 0.0;
  for (int q = lo + threadIdx.x; q < hi; q += 256) v = MIN ? fmin(v, in[q]) : v + in[q];
  sm[threadIdx.x] = v;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if (threadIdx.x < w) sm[threadIdx.x] = MIN ? fmin(sm[threadIdx.x], sm[threadIdx.x + w]) : sm[threadIdx.x] + sm[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) out[blockIdx.x] = sm[0];
}

// Both passes of k_reduce<false> in one work-group, for up to RED_FUSE
// chunks: each chunk is summed with pass 1's thread partition and tree, the
// chunk sums then with pass 2's (bit-identical to the two launches).
constexpr int RED_FUSE = 8;
__global__ __launch_bounds__(256) void k_reduce_sum1(const double* __restrict__ in, int n, double* __restrict__ out,
                                                     const int* __restrict__ halt) {
  if (halt && *halt) return;
  __shared__ double sm[256];
  __shared__ double cs[256];
  cs[threadIdx.x] = 0.0;
  const int nb = (n + RED_CHUNK - 1) / RED_CHUNK;
  for (int b = 0; b < nb; b++) {
    const int lo = b * RED_CHUNK, hi = min(n, lo + RED_CHUNK);
    double v = 0.0;
    for (int q = lo + threadIdx.x; q < hi; q += 256) v = v + in[q];
    sm[threadIdx.x] = v;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
      if (threadIdx.x < w) sm[threadIdx.x] = sm[threadIdx.x] + sm[threadIdx.x + w];
      __syncthreads();
    }
    if (threadIdx.x == 0) cs[b] = 0.0 + sm[0];  // pass 2 starts each thread's sum at 0
    __syncthreads();
  }
  for (int w = 128; w > 0; w >>= 1) {
    if (threadIdx.x < w) cs[threadIdx.x] = cs[threadIdx.x] + cs[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) *out = cs[0];
}

// Pass 2 of the dt minimum: dtglobal = min(1e5, min dt) (Mod_Explicit.f90:40,
// Mod_dt.f90:250-258) and the time advance of time-accurate runs (:42-45).
// dtglobal = min(1e5, .) (Set_Global_dt, Mod_dt.f90:86-107); the level clock
// advances on the device: time (+ dtglobal when time-accurate), iteration + 1.
__global__ __launch_bounds__(256) void k_dt_final(const double* __restrict__ in, int n, double* scal, double* clk,
                                                  int* cnt, int time_accurate) {
  if (cnt[1]) return;
  __shared__ double sm[256];
  double v = 1e300;
  for (int q = threadIdx.x; q < n; q += 256) v = fmin(v, in[q]);
  sm[threadIdx.x] = v;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if (threadIdx.x < w) sm[threadIdx.x] = fmin(sm[threadIdx.x], sm[threadIdx.x + w]);
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    const double m = fmin(1e5, sm[0]);
    const double t = clk[0];
    const double tn = time_accurate ? t + m : t;
    scal[0] = m;
    scal[1] = tn;
    clk[0] = tn;
    cnt[0] += 1;
  }
}

// Per-block sums of per-tile partials (tile lists are ordered by block):
// work-group b sums partials [rng[b], rng[b+1]) in a fixed order.
__global__ __launch_bounds__(256) void k_block_sums(const double* __restrict__ part, const int* __restrict__ rng,
                                                    double* __restrict__ out) {
  __shared__ double sm[256];
  const int b = blockIdx.x;
  double v = 0.0;
  for (int q = rng[b] + threadIdx.x; q < rng[b + 1]; q += 256) v += part[q];
  sm[threadIdx.x] = v;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if (threadIdx.x < w) sm[threadIdx.x] += sm[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) out[b] = sm[0];
}

// Level clock from the host (setup, fuss_set_clock, adjoint recompute).
__global__ void k_set_clock(double* clk, int* cnt, double time, int iter) {
  if (threadIdx.x == 0 && blockIdx.x == 0) {
    clk[0] = time;
    cnt[0] = iter;
    cnt[1] = 0;
  }
}

// fuss_run on a captured graph: counters cnt[2] (steps of this run), cnt[3]
// (steps before this chunk); a step that meets a stop test raises the halt
// flag cnt[1], which turns every later kernel of the graph into a no-op.
__global__ void k_run_begin(int* cnt, int maxcount) {
  if (threadIdx.x == 0) {
    cnt[1] = 0;
    cnt[2] = 0;
    cnt[4] = maxcount;
  }
}
__global__ void k_chunk_begin(int* cnt) {
  if (threadIdx.x == 0) cnt[3] = cnt[2];
}
__global__ void k_step_end(const double* __restrict__ scal, const double* __restrict__ clk, int* cnt,
                           const int* __restrict__ status, double* hres, double* htime, double thr, double tthr) {
  if (threadIdx.x != 0 || cnt[1]) return;
  const int maxcount = cnt[4];
  const double res = sqrt(scal[2]);
  const int n = cnt[2];
  hres[n - cnt[3]] = res;
  htime[n - cnt[3]] = clk[0];
  cnt[2] = n + 1;
  if (status[0] || n + 1 >= maxcount || res <= thr || clk[0] >= tthr) cnt[1] = 1;
}

// ------------------------------------------------------------------------
// kernels: multigrid
// ------------------------------------------------------------------------
struct MGArgs {
  BlockInfo F, Cb;
  double* fT;
  const int* fmat;
  const double* fvol;
  const double* fqvol;
  double* cT;
  int* cmat;
  const double* cvol;
  double* cqvol;
  Tab<double> tab;
  int* status;
  int bidx;
};

// fine2coarse_prim, Lib_Multigrid.f90:38-125
__global__ void k_restrict(MGArgs a) {
  const long idx = blockIdx.x * (long)blockDim.x + threadIdx.x;
  const BlockInfo& C = a.Cb;
  const BlockInfo& F = a.F;
  if (idx >= (long)C.ni * C.nj * C.nk) return;
  const int i = 1 + (int)(idx % C.ni), j = 1 + (int)((idx / C.ni) % C.nj), k = 1 + (int)(idx / ((long)C.ni * C.nj));
  const bool two_d = F.nk == 1;
  const int i2 = 2 * i, j2 = 2 * j, k2 = two_d ? 1 : 2 * k, i2d = i2 - 1, j2d = j2 - 1, k2d = two_d ? 1 : k2 - 1;
  auto fp = [&](int ii, int jj, int kk) { return F.base + (ii - 1) + (long)(jj - 1) * F.sy + (long)(kk - 1) * F.sz; };
  int bad = 0;
  auto fC = [&](int ii, int jj, int kk) {
    const long q = fp(ii, jj, kk);
    return tab_interp(a.tab.h + (long)a.fmat[q] * a.tab.nT, a.fT[q], a.tab.Ti1, a.tab.Ti2, &bad);
  };
  double cc, cq, tv;
  {
#pragma clang fp contract(off)
    const long q1 = fp(i2d, j2d, k2d), q2 = fp(i2, j2d, k2d), q3 = fp(i2d, j2, k2d), q4 = fp(i2, j2, k2d);
    if (two_d) {
      cc = fC(i2d, j2d, 1) * a.fvol[q1] + fC(i2, j2d, 1) * a.fvol[q2] + fC(i2d, j2, 1) * a.fvol[q3] +
           fC(i2, j2, 1) * a.fvol[q4];
      cq = a.fqvol[q1] * a.fvol[q1] + a.fqvol[q2] * a.fvol[q2] + a.fqvol[q3] * a.fvol[q3] + a.fqvol[q4] * a.fvol[q4];
      tv = a.fvol[q1] + a.fvol[q2] + a.fvol[q3] + a.fvol[q4];
    } else {
      const long q5 = fp(i2d, j2d, k2), q6 = fp(i2, j2d, k2), q7 = fp(i2d, j2, k2), q8 = fp(i2, j2, k2);
      cc = fC(i2d, j2d, k2d) * a.fvol[q1] + fC(i2, j2d, k2d) * a.fvol[q2] + fC(i2d, j2, k2d) * a.fvol[q3] +
           fC(i2, j2, k2d) * a.fvol[q4] + fC(i2d, j2d, k2) * a.fvol[q5] + fC(i2, j2d, k2) * a.fvol[q6] +
           fC(i2d, j2, k2) * a.fvol[q7] + fC(i2, j2, k2) * a.fvol[q8];
      cq = a.fqvol[q1] * a.fvol[q1] + a.fqvol[q2] * a.fvol[q2] + a.fqvol[q3] * a.fvol[q3] +
           a.fqvol[q4] * a.fvol[q4] + a.fqvol[q5] * a.fvol[q5] + a.fqvol[q6] * a.fvol[q6] +
           a.fqvol[q7] * a.fvol[q7] + a.fqvol[q8] * a.fvol[q8];
      tv = a.fvol[q1] + a.fvol[q2] + a.fvol[q3] + a.fvol[q4] + a.fvol[q5] + a.fvol[q6] + a.fvol[q7] +
           a.fvol[q8];
    }
  }
  const long pc = C.base + (i - 1) + (long)(j - 1) * C.sy + (long)(k - 1) * C.sz;
  a.cqvol[pc] = cq / tv;
  // cmatID = fmatID(2i,2j,2k); in 2-D the k=2 plane is the mirror of plane 1
  const int m = a.fmat[fp(i2, j2, two_d ? 1 : k2)];
  a.cmat[pc] = m;
  const double ccv = cc / a.cvol[pc];
  a.cT[pc] = tab_inverse_search(a.tab.h + (long)m * a.tab.nT, ccv, a.tab.Ti1, a.tab.Ti2, &bad);
  if (bad) flag_error(a.status, FUSS_ETABLE, a.bidx, i, j, k);
}

// coarse2fine_prim, Lib_Multigrid.f90:128-220
__global__ void k_prolong(MGArgs a) {
  const long idx = blockIdx.x * (long)blockDim.x + threadIdx.x;
  const BlockInfo& C = a.Cb;
  const BlockInfo& F = a.F;
  if (idx >= (long)C.ni * C.nj * C.nk) return;
  const int i = 1 + (int)(idx % C.ni), j = 1 + (int)((idx / C.ni) % C.nj), k = 1 + (int)(idx / ((long)C.ni * C.nj));
  const bool two_d = F.nk == 1;
  double co[8];
  if (two_d) {
    const double a1 = 9.0 / 16.0, a2 = 3.0 / 16.0, a3 = 1.0 / 16.0, a4 = 0.0;
    co[0] = a1; co[1] = a2; co[2] = a2; co[3] = a4; co[4] = a3; co[5] = a4; co[6] = a4; co[7] = a4;
  } else {
    const double a1 = 27.0 / 64.0, a2 = 9.0 / 64.0, a3 = 3.0 / 64.0, a4 = 1.0 / 64.0;
    co[0] = a1; co[1] = a2; co[2] = a2; co[3] = a2; co[4] = a3; co[5] = a3; co[6] = a3; co[7] = a4;
  }
  const int id[6] = {i - 1 > 1 ? i - 1 : 1, j - 1 > 1 ? j - 1 : 1, k - 1 > 1 ? k - 1 : 1,
                     i + 1 < C.ni ? i + 1 : C.ni, j + 1 < C.nj ? j + 1 : C.nj, k + 1 < C.nk ? k + 1 : C.nk};
  auto cP = [&](int ii, int jj, int kk) { return a.cT[C.base + (ii - 1) + (long)(jj - 1) * C.sy + (long)(kk - 1) * C.sz]; };
  const int i2 = 2 * i, j2 = 2 * j, k2 = two_d ? 1 : 2 * k;
  const int i2d = i2 - 1, j2d = j2 - 1, k2d = two_d ? 1 : k2 - 1;
  int counter = 0;
  for (int kk = k2d; kk <= k2; kk++)
    for (int jj = j2d; jj <= j2; jj++)
      for (int ii = i2d; ii <= i2; ii++) {
        const int m0 = (counter & 1) ? 3 : 0, m1 = (counter & 2) ? 4 : 1, m2 = (counter & 4) ? 5 : 2;
        const int A0 = id[m0], B1 = id[m1], C2 = id[m2];
        double v;
        {
#pragma clang fp contract(off)
          v = co[0] * cP(i, j, k) + co[1] * cP(A0, j, k) + co[2] * cP(i, B1, k) + co[3] * cP(i, j, C2) +
              co[4] * cP(A0, B1, k) + co[5] * cP(A0, j, C2) + co[6] * cP(i, B1, C2) + co[7] * cP(A0, B1, C2);
        }
        a.fT[F.base + (ii - 1) + (long)(jj - 1) * F.sy + (long)(kk - 1) * F.sz] = v;
        counter++;
      }
}

#include "fuss_refine.h"
#include "fuss_adjoint.h"

// ------------------------------------------------------------------------
// host side
// ------------------------------------------------------------------------

namespace {

SerDev serdev(fuss_ctx* c) {
  SerDev d;
  d.t = c->ser_t.p; d.v = c->ser_v.p; d.off = c->ser_off.p; d.n = c->ser_n.p; d.per = c->ser_per.p;
  return d;
}

long cidx(const BlockInfo& B, int i, int j, int k) {
  return B.base + (i - 1) + (long)(j - 1) * B.sy + (long)(k - 1) * B.sz;
}

// Compute_Norm_Area sign per direction (Lib_Metrics.f90:86-153), host side.
double norm_sign(const HostBlock& h, int dir) {
  auto N = [&](int i, int j, int k) { return &h.nodes[3 * (i + (long)(h.n[0] + 1) * (j + (long)(h.n[1] + 1) * k))]; };
  double d1[3], d2[3], sn[3], c1[3], c2[3];
  const double *a, *b, *c, *d;
  if (dir == 0) { for (int x = 0; x < 3; x++) { d1[x] = N(1, 1, 1)[x] - N(1, 0, 0)[x]; d2[x] = N(1, 0, 1)[x] - N(1, 1, 0)[x]; } }
  else if (dir == 1) { for (int x = 0; x < 3; x++) { d1[x] = N(0, 1, 1)[x] - N(1, 1, 0)[x]; d2[x] = N(1, 1, 1)[x] - N(0, 1, 0)[x]; } }
  else { for (int x = 0; x < 3; x++) { d1[x] = N(0, 1, 1)[x] - N(1, 0, 1)[x]; d2[x] = N(0, 0, 1)[x] - N(1, 1, 1)[x]; } }
  sn[0] = d1[1] * d2[2] - d1[2] * d2[1];
  sn[1] = d1[2] * d2[0] - d1[0] * d2[2];
  sn[2] = d1[0] * d2[1] - d1[1] * d2[0];
  if (dir == 0) { a = N(1, 1, 1); b = N(1, 0, 1); c = N(1, 0, 0); d = N(1, 1, 0); }
  else if (dir == 1) { a = N(1, 1, 1); b = N(0, 1, 1); c = N(0, 1, 0); d = N(1, 1, 0); }
  else { a = N(1, 1, 1); b = N(0, 1, 1); c = N(0, 0, 1); d = N(1, 0, 1); }
  for (int x = 0; x < 3; x++) c1[x] = 0.25 * (a[x] + b[x] + c[x] + d[x]);
  if (dir == 0) { a = N(0, 1, 1); b = N(0, 0, 1); c = N(0, 0, 0); d = N(0, 1, 0); }
  else if (dir == 1) { a = N(1, 0, 1); b = N(0, 0, 1); c = N(0, 0, 0); d = N(1, 0, 0); }
  else { a = N(1, 1, 0); b = N(0, 1, 0); c = N(0, 0, 0); d = N(1, 0, 0); }
  for (int x = 0; x < 3; x++) c2[x] = 0.25 * (a[x] + b[x] + c[x] + d[x]);
  const double scal = (c1[0] - c2[0]) * sn[0] + (c1[1] - c2[1]) * sn[1] + (c1[2] - c2[2]) * sn[2];
  return scal >= 0.0 ? 1.0 : -1.0;
}

// Largest relative deviation of the per-axis factors from the stored
// volumes and normal weights for which a level uses them (about 4 ulp of
// rounding-order difference is expected; anything larger keeps the stored
// arrays).
constexpr double RECT_TOL = 1e-14;

// Per-axis metric factors (RectAxes) of a rectilinear block, from the same
// Compute_Metric_Tensor / Compute_Norm_Area arithmetic as the 3-D arrays
// (cells and faces on the block's first line in each direction).  Returns
// false when some node is off the block's coordinate planes.
bool rect_factors(const HostBlock& h, std::vector<double>& out) {
  const int n[3] = {h.n[0], h.n[1], h.n[2]};
  auto N = [&](int i, int j, int k) { return &h.nodes[3 * (i + (size_t)(n[0] + 1) * (j + (size_t)(n[1] + 1) * k))]; };
  for (int k = 0; k <= n[2]; k++)
    for (int j = 0; j <= n[1]; j++)
      for (int i = 0; i <= n[0]; i++) {
        const double* v = N(i, j, k);
        if (v[0] != N(i, 0, 0)[0] || v[1] != N(0, j, 0)[1] || v[2] != N(0, 0, k)[2]) return false;
      }
  for (int d = 0; d < 3; d++) {
    const int nd = n[d];
    std::vector<double> a(nd + 2), m(nd + 2), f(nd + 2, 0.0), q(nd + 2);
    for (int s = 1; s <= nd; s++) {
      int c[3] = {1, 1, 1};
      c[d] = s;
      const int i = c[0], j = c[1], k = c[2];
      double M[9], dl[3], v;
      if (metric_tensor(N(i - 1, j - 1, k - 1), N(i - 1, j - 1, k), N(i - 1, j, k - 1), N(i - 1, j, k),
                        N(i, j - 1, k - 1), N(i, j - 1, k), N(i, j, k - 1), N(i, j, k), M, dl, &v))
        return false;
      // edge along d: the node difference the face areas and tetrahedra use
      // (dl_d sums four such differences in coordinate order, which rounds
      // at the coordinates' ulp rather than the edge's)
      a[s] = N(i, j, k)[d] - N(c[0] - (d == 0), c[1] - (d == 1), c[2] - (d == 2))[d];
      m[s] = M[4 * d];  // M_dd
      q[s] = 1.0 / (dl[d] * dl[d]);
    }
    a[0] = a[1]; a[nd + 1] = a[nd];
    q[0] = q[1]; q[nd + 1] = q[nd];
    // the normal's sign on the direction's faces: face_normal of the first
    // face as k_face_metrics evaluates it (its component d is +-1 exactly)
    const double* p4[4];
    if (d == 0) { p4[0] = N(1, 1, 1); p4[1] = N(1, 0, 0); p4[2] = N(1, 0, 1); p4[3] = N(1, 1, 0); }
    else if (d == 1) { p4[0] = N(0, 1, 1); p4[1] = N(1, 1, 0); p4[2] = N(1, 1, 1); p4[3] = N(0, 1, 0); }
    else { p4[0] = N(0, 1, 1); p4[1] = N(1, 0, 1); p4[2] = N(0, 0, 1); p4[3] = N(1, 1, 1); }
    double Nv[3], A;
    face_normal(p4[0], p4[1], p4[2], p4[3], norm_sign(h, d), Nv, &A);
    if (!(Nv[d] == 1.0 || Nv[d] == -1.0)) return false;
    for (int s = 1; s < nd; s++) {
#pragma clang fp contract(off)
      f[s] = 0.5 * (m[s] + m[s + 1]) * Nv[d];
    }
    out.insert(out.end(), a.begin(), a.end());
    out.insert(out.end(), f.begin(), f.end());
    out.insert(out.end(), q.begin(), q.end());
  }
  return true;
}

int upload(void* dst, const void* src, size_t bytes);

// Per-axis factors of a 3-D level whose blocks are all rectilinear, checked
// on the device against the stored metrics (after k_cell_metrics and
// k_face_metrics, with L.hblk uploaded); sets L.rect_w / L.rect_dt and the
// blocks' roff (-1 again when neither check passes).
int rect_prepare(fuss_ctx* c, Level& L) {
  L.rect_w = L.rect_dt = 0;
  L.rect_dev = -1.0;
  if (c->ndir != 3 || L.kp != GC || L.tang || !L.dims2) return 0;
  std::vector<double> fac;
  for (int b = 0; b < L.nb; b++) {
    const size_t o = fac.size();
    if (!rect_factors(L.hb[b], fac)) {
      for (auto& B : L.hblk) B.roff = -1;
      return 0;
    }
    L.hblk[b].roff = (int)o;
  }
  if (L.rfac.alloc(fac.size())) return FUSS_EHIP;
  int e = upload(L.rfac.p, fac.data(), fac.size() * 8);
  e |= upload(L.blk.p, L.hblk.data(), sizeof(BlockInfo) * L.nb);
  DBuf<unsigned long long> dev;
  DBuf<int> nbad;
  if (e || dev.alloc(1) || nbad.alloc(1)) return FUSS_EHIP;
  HIPCHK(hipMemsetAsync(dev.p, 0, sizeof(unsigned long long), c->stream));
  HIPCHK(hipMemsetAsync(nbad.p, 0, sizeof(int), c->stream));
  k_rect_check<<<(unsigned)L.htiles.size(), TILE_T, 0, c->stream>>>(L.blk.p, L.tiles.p, L.rfac.p, L.vol.p, L.invL2.p,
                                                                    L.W[0].p, L.W[4].p, L.W[8].p, dev.p, nbad.p);
  HIPCHK(hipGetLastError());
  unsigned long long hd = 0;
  int hb = 0;
  HIPCHK(hipMemcpyAsync(&hd, dev.p, sizeof(hd), hipMemcpyDeviceToHost, c->stream));
  HIPCHK(hipMemcpyAsync(&hb, nbad.p, sizeof(hb), hipMemcpyDeviceToHost, c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  dev.free();
  nbad.free();
  double dv;
  memcpy(&dv, &hd, sizeof(dv));
  L.rect_dev = dv;
  L.rect_w = dv <= RECT_TOL;
  L.rect_dt = hb == 0;
  if (!L.rect_w && !L.rect_dt) {
    for (auto& B : L.hblk) B.roff = -1;
    L.rfac.free();
    if (upload(L.blk.p, L.hblk.data(), sizeof(BlockInfo) * L.nb)) return FUSS_EHIP;
  }
  return 0;
}

// BC_Connect_Metrics, Lib_Metrics.f90:704-936: Mg(1) of a 101 entry from the
// local face nodes and the donor's nodes (host side, surface only).
int connect_metric(const HostBlock& bm, const HostBlock& bs, const int* q, double* Mg1) {
  const int e = connect_metric_nodes(bm.nodes.data(), bm.n[0], bm.n[1], bm.n[2], bs.nodes.data(), bs.n[0], bs.n[1],
                                     bs.n[2], q, Mg1);
  return e == 1 ? FUSS_EBC : (e ? FUSS_EMETRIC : 0);
}

// BC_Extrapolate_Metrics, Lib_Metrics.f90:379-600, first ghost layer: the
// ghost cell's nodes extrapolated along the face normal from the block's node
// layers -- cubic (3a - 3b + c), linear (2a - b) across j faces of blocks with
// 2 or fewer j cells and across k faces of 2-D levels -- and their metric
// tensor Mg(1) (host side, chimera entries).
int extrap_metric(const HostBlock& h, int Im, int Jm, int Km, int Fm, int ndir, double* Mg1) {
  const bool cubic = !((Fm == 3 || Fm == 4) && h.n[1] <= 2) && !((Fm == 5 || Fm == 6) && ndir != 3);
  int base[4][3], di = 0, dj = 0, dk = 0;
  switch (Fm) {
    case 1: { int q[4][3] = {{Im - 1, Jm - 1, Km - 1}, {Im - 1, Jm - 1, Km}, {Im - 1, Jm, Km - 1}, {Im - 1, Jm, Km}}; memcpy(base, q, sizeof(base)); di = 1; } break;
    case 2: { int q[4][3] = {{Im, Jm - 1, Km - 1}, {Im, Jm - 1, Km}, {Im, Jm, Km - 1}, {Im, Jm, Km}}; memcpy(base, q, sizeof(base)); di = -1; } break;
    case 3: { int q[4][3] = {{Im - 1, Jm - 1, Km - 1}, {Im - 1, Jm - 1, Km}, {Im, Jm - 1, Km - 1}, {Im, Jm - 1, Km}}; memcpy(base, q, sizeof(base)); dj = 1; } break;
    case 4: { int q[4][3] = {{Im - 1, Jm, Km - 1}, {Im - 1, Jm, Km}, {Im, Jm, Km - 1}, {Im, Jm, Km}}; memcpy(base, q, sizeof(base)); dj = -1; } break;
    case 5: { int q[4][3] = {{Im - 1, Jm - 1, Km - 1}, {Im - 1, Jm, Km - 1}, {Im, Jm - 1, Km - 1}, {Im, Jm, Km - 1}}; memcpy(base, q, sizeof(base)); dk = 1; } break;
    case 6: { int q[4][3] = {{Im - 1, Jm - 1, Km}, {Im - 1, Jm, Km}, {Im, Jm - 1, Km}, {Im, Jm, Km}}; memcpy(base, q, sizeof(base)); dk = -1; } break;
    default: return FUSS_EBC;
  }
  double L0[4][3], G[4][3];
  for (int q = 0; q < 4; q++) {
    const double* p[3];
    for (int s2 = 0; s2 <= (cubic ? 2 : 1); s2++) {
      const int ii = base[q][0] + di * s2, jj = base[q][1] + dj * s2, kk = base[q][2] + dk * s2;
      if (ii < 0 || ii > h.n[0] || jj < 0 || jj > h.n[1] || kk < 0 || kk > h.n[2]) return FUSS_EBC;
      p[s2] = &h.nodes[3 * (ii + (long)(h.n[0] + 1) * (jj + (long)(h.n[1] + 1) * kk))];
    }
    {
#pragma clang fp contract(off)
      for (int x = 0; x < 3; x++) {
        L0[q][x] = p[0][x];
        G[q][x] = cubic ? 3.0 * p[0][x] - 3.0 * p[1][x] + p[2][x] : 2.0 * p[0][x] - p[1][x];
      }
    }
  }
  double dl[3], v;
  int e = 0;
  switch (Fm) {
    case 1: e = metric_tensor(G[0], G[1], G[2], G[3], L0[0], L0[1], L0[2], L0[3], Mg1, dl, &v); break;
    case 2: e = metric_tensor(L0[0], L0[1], L0[2], L0[3], G[0], G[1], G[2], G[3], Mg1, dl, &v); break;
    case 3: e = metric_tensor(G[0], G[1], L0[0], L0[1], G[2], G[3], L0[2], L0[3], Mg1, dl, &v); break;
    case 4: e = metric_tensor(L0[0], L0[1], G[0], G[1], L0[2], L0[3], G[2], G[3], Mg1, dl, &v); break;
    case 5: e = metric_tensor(G[0], L0[0], G[1], L0[1], G[2], L0[2], G[3], L0[3], Mg1, dl, &v); break;
    default: e = metric_tensor(L0[0], G[0], L0[1], G[1], L0[2], G[2], L0[3], G[3], Mg1, dl, &v); break;
  }
  return e ? FUSS_EMETRIC : 0;
}

int upload(void* dst, const void* src, size_t bytes) {
  return hipMemcpy(dst, src, bytes, hipMemcpyHostToDevice) == hipSuccess ? 0 : FUSS_EHIP;
}

// The state (or the tables) changed outside the step loop: the next step
// computes its dt afresh.
void drop_dt_next(fuss_ctx* c) {
  for (auto& L : c->lev) L.dt_next = 0;
}

void run_graph_free(Level& L) {
  if (L.run_exec) (void)hipGraphExecDestroy(L.run_exec);
  if (L.run_graph) (void)hipGraphDestroy(L.run_graph);
  L.run_exec = nullptr;
  L.run_graph = nullptr;
  L.run_bufs.clear();
}

// Every device buffer a captured step of level L may address (kernel
// arguments are fixed at capture): a replay is allowed only while all of them
// are the allocations the graph was captured with.
std::vector<const void*> run_buffers(const fuss_ctx* c, const Level& L) {
  std::vector<const void*> v = {
      L.blk.p, L.tiles.p, L.ctiles.p, L.c8tiles.p, L.xytiles.p, L.Tb[0].p, L.Tb[1].p, L.Tb[2].p, L.R.p, L.R2.p,
      L.dt.p, L.qvol.p, L.vol.p, L.invL2.p, L.mat.p, L.bc.p, L.bch.p, L.gmap.p, L.fmap.p, L.bcv.p, L.bcf.p,
      L.rfac.p, L.ttx.p, L.tty.p, L.ttz.p, L.tri.p, L.trioff.p, L.chr.p, L.chd.p, L.chv.p, L.clk.p, L.cnt.p,
      L.partials.p, L.pmin.p, L.red2.p, c->dcp.p, c->drho.p, c->dkap.p, c->dh.p, c->dkap2.p, c->dpr4.p,
      c->status.p, c->scal.p, c->run_hist.p, c->ser_t.p, c->ser_v.p, c->ser_off.p, c->ser_n.p, c->ser_per.p,
      L.cd.p, L.ci.p};
  for (const auto& w : L.W) v.push_back(w.p);
  for (const auto& r : L.rng) v.push_back(r.p);
  const auto& F = L.f;
  for (const auto* q : {F.Tb[0].p, F.Tb[1].p, F.Tb[2].p, F.R.p, F.R2.p, F.dt.p, F.qvol.p, F.vol.p, F.bcv.p, F.bcf.p, F.W[0].p, F.W[1].p, F.W[2].p, F.W[3].p, F.W[4].p, F.W[5].p, F.W[6].p,
                        F.W[7].p, F.W[8].p})
    v.push_back(q);
  for (const auto* q : {c->fcp.p, c->frho.p, c->fkap.p, c->fh.p, c->fkap2.p, c->fpr4.p}) v.push_back(q);
  return v;
}

// hrng[q]: index of the first tile of each block in a block-ordered list
void block_ranges(Level& L, int q, const std::vector<Tile>& list) {
  L.hrng[q].assign(L.nb + 1, (int)list.size());
  for (int t = (int)list.size() - 1; t >= 0; t--) L.hrng[q][list[t].b] = t;
  for (int b = L.nb - 1; b >= 0; b--) L.hrng[q][b] = std::min(L.hrng[q][b], L.hrng[q][b + 1]);
}

// Column tiles of kc planes on 3-D levels: 64x4 (stage passes, z pass) and
// 64x8 (fused x-y-z IRS pass), with their block ranges.
void column_tiles(Level& L) {
  L.hctiles.clear();
  L.hc8tiles.clear();
  if (L.kp == GC)
    for (int b = 0; b < L.nb; b++) {
      const BlockInfo& B = L.hblk[b];
      for (int k = 1; k <= B.nk; k += L.kc) {
        for (int j = 1; j <= B.nj; j += TILE_Y)
          for (int i = 1; i <= B.ni; i += TILE_X) L.hctiles.push_back(Tile{b, i, j, k});
        for (int j = 1; j <= B.nj; j += 2 * TILE_Y)
          for (int i = 1; i <= B.ni; i += TILE_X) L.hc8tiles.push_back(Tile{b, i, j, k});
      }
    }
  block_ranges(L, 1, L.hctiles);
  block_ranges(L, 3, L.hc8tiles);
}

int build_level_layout(fuss_ctx* c, Level& L) {
  long off = 0, noff = 0;
  L.hblk.resize(L.nb);
  L.htiles.clear();
  L.hxytiles.clear();
  for (int b = 0; b < L.nb; b++) {
    BlockInfo& B = L.hblk[b];
    const HostBlock& h = L.hb[b];
    B.ni = h.n[0]; B.nj = h.n[1]; B.nk = h.n[2];
    // rows of whole 128-byte lines, and cell (1,1,1) on a line boundary: every
    // interior row then starts a line, so a 64-cell tile row is 4 full lines
    // (unaligned rows straddle 5, the extra line fetched by two work-groups)
    B.sy = (B.ni + 2 * GC + ROW_ALIGN - 1) / ROW_ALIGN * ROW_ALIGN;
    B.sz = (long)B.sy * (B.nj + 2 * GC);
    off = (off + GC + ROW_ALIGN - 1) / ROW_ALIGN * ROW_ALIGN - GC;
    B.off = off;
    B.npad = B.sz * (B.nk + 2 * L.kp);
    B.base = off + GC + (long)GC * B.sy + (long)L.kp * B.sz;
    B.node_off = noff;
    for (int f = 0; f < 6; f++) B.fmap_off[f] = -1;
    B.face_active = 0;
    B.conn = 0;
    B.uqv = 0;
    B.roff = -1;
    B.tang = 0;
    B.umat = -1;
    off += B.npad;
    noff += (long)(B.ni + 1) * (B.nj + 1) * (B.nk + 1);
    for (int k = 1; k <= B.nk; k++)
      for (int j = 1; j <= B.nj; j += TILE_Y)
        for (int i = 1; i <= B.ni; i += TILE_X) L.htiles.push_back(Tile{b, i, j, k});
    for (int k = 1; k <= B.nk; k++)
      for (int j = 1; j <= B.nj; j += XY_TJ)
        for (int i = 1; i <= B.ni; i += TILE_X) L.hxytiles.push_back(Tile{b, i, j, k});
    L.ncell += (long)B.ni * B.nj * B.nk;
  }
  L.npad = off;
  L.nnodes = noff;
  L.dims2 = 1;
  for (int b = 0; b < L.nb; b++) L.dims2 &= L.hblk[b].ni >= 2 && L.hblk[b].nj >= 2 && L.hblk[b].nk >= 2;
  block_ranges(L, 0, L.htiles);
  block_ranges(L, 2, L.hxytiles);
  column_tiles(L);
  (void)c;
  // kernels address a level's arrays with 32-bit byte offsets
  if (L.npad * 8 >= (1L << 32)) return FUSS_EUNSUPPORTED;
  return 0;
}

int alloc_level(fuss_ctx* c, Level& L) {
  int e = 0;
  for (int q = 0; q < 3; q++) e |= L.Tb[q].alloc(L.npad);
  e |= L.R.alloc(L.npad);
  e |= L.R2.alloc(L.npad);
  e |= L.dt.alloc(L.npad);
  e |= L.qvol.alloc(L.npad);
  e |= L.vol.alloc(L.npad);
  e |= L.invL2.alloc(L.npad);
  e |= L.mat.alloc(L.npad);
  e |= L.M.alloc(9 * L.npad);
  for (int d = 0; d < 3; d++) e |= L.F[d].alloc(4 * L.npad);
  for (int w = 0; w < 9; w++) e |= L.W[w].alloc(L.npad);
  e |= L.nodes.alloc(3 * L.nnodes);
  e |= L.tiles.alloc(L.htiles.size());
  e |= L.ctiles.alloc(std::max<size_t>(L.hctiles.size(), 1));
  e |= L.xytiles.alloc(std::max<size_t>(L.hxytiles.size(), 1));
  e |= L.c8tiles.alloc(std::max<size_t>(L.hc8tiles.size(), 1));
  const size_t np = std::max<size_t>(std::max(L.htiles.size(), L.hctiles.size()), 1);
  e |= L.partials.alloc(np);
  e |= L.pmin.alloc(np);
  e |= L.clk.alloc(2);
  e |= L.cnt.alloc(5);
  for (int q = 0; q < 4; q++) {
    e |= L.rng[q].alloc(L.nb + 1);
    if (!e) e |= upload(L.rng[q].p, L.hrng[q].data(), (L.nb + 1) * sizeof(int));
  }
  e |= L.red2.alloc(np / RED_CHUNK + 2);
  (void)c;
  return e ? FUSS_EHIP : 0;
}

int build_bc(fuss_ctx* c, Level& L) {
  const int n = L.nbound;
  L.no_ycon = 1;
  L.no_zcon = 1;
  L.no_xcon = 1;
  L.ghost_read = 0;
  // grouping check: (block, face) non-decreasing, as BC_Fluxes assumes
  // (Mod_BC_Fluxes.f90:26-33)
  for (int e = 1; e < n; e++) {
    const int* a = &L.hib[15 * (e - 1)];
    const int* b = &L.hib[15 * e];
    if (b[0] < a[0] || (b[0] == a[0] && b[4] < a[4])) return FUSS_EBC;
  }
  // face maps
  std::vector<int> fmap;
  for (int b = 0; b < L.nb; b++) {
    BlockInfo& B = L.hblk[b];
    for (int f = 1; f <= 6; f++) {
      const int ext = f <= 2 ? B.nj * B.nk : (f <= 4 ? B.ni * B.nk : B.ni * B.nj);
      B.fmap_off[f - 1] = (int)fmap.size();
      fmap.insert(fmap.end(), ext, -1);
    }
  }
  std::vector<BCRec> recs(n);
  for (int e = 0; e < n; e++) {
    const int* q = &L.hib[15 * e];
    const double* r = &L.hrb[6 * e];
    BCRec& R = recs[e];
    memset(&R, 0, sizeof(R));
    const int b = q[0] - 1, i = q[1], j = q[2], k = q[3], f = q[4], type = q[5];
    if (b < 0 || b >= L.nb || f < 1 || f > 6) return FUSS_EBC;
    BlockInfo& B = L.hblk[b];
    if (i < 1 || i > B.ni || j < 1 || j > B.nj || k < 1 || k > B.nk) return FUSS_EBC;
    const bool on = (f == 1 && i == 1) || (f == 2 && i == B.ni) || (f == 3 && j == 1) || (f == 4 && j == B.nj) ||
                    (f == 5 && k == 1) || (f == 6 && k == B.nk);
    if (!on) return FUSS_EBC;
    const int mi = f <= 2 ? (j - 1) + B.nj * (k - 1) : (f <= 4 ? (i - 1) + B.ni * (k - 1) : (i - 1) + B.ni * (j - 1));
    if (fmap[B.fmap_off[f - 1] + mi] != -1) return FUSS_EBC;
    fmap[B.fmap_off[f - 1] + mi] = e;
    const int fpos = B.fmap_off[f - 1] + mi;
    if (type == 102 && &L != &c->lev[0]) return FUSS_EUNSUPPORTED;  // donor lists: finest level only
    R.type = type; R.f = f; R.dir = (f + 1) / 2; R.series = L.hser.empty() ? -1 : L.hser[e];
    R.b = b; R.i = i; R.j = j; R.k = k;
    R.cell = cidx(B, i, j, k);
    R.fpos = fpos;
    R.qw = r[0]; R.Tw = r[1]; R.hconv = r[2]; R.Tref = r[4]; R.eps = r[5];
    const int gi = guide(f, 0), gj = guide(f, 1), gk = guide(f, 2);
    // layer-1 ghost (2-D levels: k faces address the single k plane kept on
    // each side of the block)
    R.gdst = cidx(B, i - gi, j - gj, k - gk);
    R.gsrc = R.cell;  // Ghost_Symmetry layer 1: T(ghost) = T(cell)
    R.rdonor = -1;
    for (int t = 0; t < 4; t++) R.dn[t] = R.cell;
    if (type == 101) {
      const int bs = q[6] - 1, Is = q[7], Js = q[8], Ks = q[9], Fs = q[10];
      if (bs < 0 || bs >= L.nb || Fs < 1 || Fs > 6) return FUSS_EBC;
      const BlockInfo& S = L.hblk[bs];
      if (Is < 1 || Is > S.ni || Js < 1 || Js > S.nj || Ks < 1 || Ks > S.nk) return FUSS_EBC;
      R.gsrc = cidx(S, Is, Js, Ks);  // Ghost_Connection g=1: donor (Is,Js,Ks)
      R.rdonor = R.gsrc;
      const int d11 = q[11], d12 = q[12], d21 = q[13], d22 = q[14];
      int p[4][3];
      if (Fs <= 2) { int t[4][3] = {{Is, Js - d11, Ks - d12}, {Is, Js + d11, Ks + d12}, {Is, Js - d21, Ks - d22}, {Is, Js + d21, Ks + d22}}; memcpy(p, t, sizeof(p)); }
      else if (Fs <= 4) { int t[4][3] = {{Is - d11, Js, Ks - d12}, {Is + d11, Js, Ks + d12}, {Is - d21, Js, Ks - d22}, {Is + d21, Js, Ks + d22}}; memcpy(p, t, sizeof(p)); }
      else { int t[4][3] = {{Is - d11, Js - d12, Ks}, {Is + d11, Js + d12, Ks}, {Is - d21, Js - d22, Ks}, {Is + d21, Js + d22, Ks}}; memcpy(p, t, sizeof(p)); }
      for (int t = 0; t < 4; t++) {
        // (2-D levels keep one k plane on each side, so k offsets of +-1 stay
        // inside the block's arena box; those Tg(5:6) terms enter multiplied
        // by a vanishing M_con(3,:).N)
        if (p[t][0] < 0 || p[t][0] > S.ni + 1 || p[t][1] < 0 || p[t][1] > S.nj + 1 || p[t][2] < 0 ||
            p[t][2] > S.nk + 1)
          return FUSS_EBC;
        R.dn[t] = cidx(S, p[t][0], p[t][1], p[t][2]);
      }
      const int e2 = connect_metric(L.hb[b], L.hb[bs], q, R.Mg);
      if (e2) return e2;
      // matIDg: the donor cell's material (Ghost_Connection copies it)
    }
    if (type == 102) {
      // chimera: the flux reads its ghost (filled by k_chimera) as Tg(1),
      // the boundary cell's own tangential neighbours as Tg(3:6)
      // (Fill_BC_Ghost_Chimera, Lib_Ghost.f90:354-433) and the extrapolated
      // ghost metric (BC_Extrapolate_Metrics, Lib_Metrics.f90:379-600)
      R.gsrc = R.gdst;
      int p[4][3];
      if (f <= 2) { int t[4][3] = {{i, j - 1, k}, {i, j + 1, k}, {i, j, k - 1}, {i, j, k + 1}}; memcpy(p, t, sizeof(p)); }
      else if (f <= 4) { int t[4][3] = {{i - 1, j, k}, {i + 1, j, k}, {i, j, k - 1}, {i, j, k + 1}}; memcpy(p, t, sizeof(p)); }
      else { int t[4][3] = {{i - 1, j, k}, {i + 1, j, k}, {i, j - 1, k}, {i, j + 1, k}}; memcpy(p, t, sizeof(p)); }
      for (int t = 0; t < 4; t++) R.dn[t] = cidx(B, p[t][0], p[t][1], p[t][2]);
      const int e2 = extrap_metric(L.hb[b], i, j, k, f, c->ndir, R.Mg);
      if (e2) return e2;
    }
    if (type == 101 && (f == 3 || f == 4)) L.no_ycon = 0;
    if (type == 101 && (f == 5 || f == 6)) L.no_zcon = 0;
    if (type == 101 && (f == 1 || f == 2)) L.no_xcon = 0;
    if (type == 101 || type == 102 || type == 300 || type == 302) L.ghost_read = 1;
    const bool active = type == 101 || type == 102 || (type >= 300 && type <= 305);
    if (active) B.face_active |= 1 << (f - 1);
    if (type == 101) B.conn |= 1 << (f - 1);
  }
  for (size_t q = 0; q < fmap.size(); q++)
    if (fmap[q] < 0) return FUSS_EBC;
  L.any_active = 0;
  for (int b = 0; b < L.nb; b++) L.any_active |= L.hblk[b].face_active;
  // (ghost, source) pairs of the ghost fill, 32-bit arena indices
  std::vector<int2> gm(std::max(n, 1));
  for (int e = 0; e < n; e++)
    gm[e] = make_int2(recs[e].gdst >= 0 && recs[e].type != 102 ? (int)recs[e].gdst : -1, (int)recs[e].gsrc);
  // chimera entries: every 102 entry needs its donor list (interior cells)
  std::vector<int4> chr;
  std::vector<int> chd;
  std::vector<double> chv;
  {
    std::vector<int> slot(n, -1);
    size_t off = 0;
    std::vector<size_t> offs;
    for (size_t q = 0; q < L.hch_ent.size(); q++) {
      const int ent = L.hch_ent[q];
      if (ent < 0 || ent >= n || recs[ent].type != 102 || slot[ent] >= 0) return FUSS_EBC;
      slot[ent] = (int)q;
      offs.push_back(off);
      off += (size_t)(L.hch_n1[q] + L.hch_n2[q]);
    }
    for (int e = 0; e < n; e++) {
      if (recs[e].type != 102) continue;
      const int q = slot[e];
      if (q < 0 || L.hch_n1[q] < 1) return FUSS_EBC;
      const int first = (int)chd.size();
      int mat = 0;
      for (int t = 0; t < L.hch_n1[q]; t++) {  // the first ghost layer's donors
        const int* dn = &L.hch_don[4 * (offs[q] + t)];
        const int bs = dn[0] - 1;
        if (bs < 0 || bs >= L.nb) return FUSS_EBC;
        const BlockInfo& S = L.hblk[bs];
        if (dn[1] < 1 || dn[1] > S.ni || dn[2] < 1 || dn[2] > S.nj || dn[3] < 1 || dn[3] > S.nk) return FUSS_EBC;
        chd.push_back((int)cidx(S, dn[1], dn[2], dn[3]));
        chv.push_back(L.hch_vf[offs[q] + t]);
        const HostBlock& H = L.hb[bs];
        mat = (int)H.matID[(dn[1] - 1) + (size_t)H.n[0] * ((dn[2] - 1) + (size_t)H.n[1] * (dn[3] - 1))] - 1;
      }
      // Ghost_Chimera leaves the last donor's material in the ghost (matIDg)
      recs[e].matg = mat;
      chr.push_back(make_int4((int)recs[e].gdst, first, L.hch_n1[q], mat));
    }
  }
  L.nchim = (int)chr.size();
  int e = L.bc.alloc(std::max(n, 1));
  e |= L.bch.alloc(std::max(n, 1));
  e |= L.gmap.alloc(gm.size());
  e |= L.bcv.alloc(std::max(n, 1));
  e |= L.bcf.alloc(fmap.size() + 1);  // (+ an entry that stays 0)
  e |= L.fmap.alloc(std::max<size_t>(fmap.size(), 1));
  if (e) return FUSS_EHIP;
  if (n) e |= upload(L.bc.p, recs.data(), sizeof(BCRec) * n);
  if (n) e |= upload(L.gmap.p, gm.data(), sizeof(int2) * n);
  if (L.nchim) {
    if (L.chr.alloc(chr.size()) || L.chd.alloc(chd.size()) || L.chv.alloc(chv.size())) return FUSS_EHIP;
    e |= upload(L.chr.p, chr.data(), sizeof(int4) * chr.size());
    e |= upload(L.chd.p, chd.data(), sizeof(int) * chd.size());
    e |= upload(L.chv.p, chv.data(), sizeof(double) * chv.size());
  }
  if (!fmap.empty()) e |= upload(L.fmap.p, fmap.data(), sizeof(int) * fmap.size());
  (void)c;
  return e;
}

// Face-map offsets of every block: faces 1..6 of block 1, then block 2, ...
// (one slot per boundary face cell); returns the face map's length.
size_t face_map_offsets(Level& L) {
  size_t nf = 0;
  for (auto& B : L.hblk)
    for (int f = 1; f <= 6; f++) {
      B.fmap_off[f - 1] = (int)nf;
      nf += (size_t)(f <= 2 ? B.nj * B.nk : (f <= 4 ? B.ni * B.nk : B.ni * B.nj));
    }
  return nf;
}

bool level_has_chimera(const Level& L) {
  for (int e = 0; e < L.nbound; e++)
    if (L.hib[15 * (size_t)e + 5] == 102) return true;
  return !L.hch_ent.empty();
}

// build_bc's records, ghost pairs and face map formed on the device
// (k_build_bc) for a level without chimera entries; needs the level's nodes
// and block infos with the face-map offsets of face_map_offsets on the device.
// Merges the face activity and connection bits into L.hblk.
int build_bc_dev(fuss_ctx* c, Level& L, size_t nf) {
  const int n = L.nbound;
  L.no_ycon = L.no_zcon = L.no_xcon = 1;
  L.ghost_read = 0;
  L.nchim = 0;
  int e = L.bc.alloc(std::max(n, 1));
  e |= L.bch.alloc(std::max(n, 1));
  e |= L.gmap.alloc(std::max(n, 1));
  e |= L.bcv.alloc(std::max(n, 1));
  e |= L.bcf.alloc(nf + 1);  // (+ an entry that stays 0)
  e |= L.fmap.alloc(std::max<size_t>(nf, 1));
  if (e) return FUSS_EHIP;
  DBuf<int> dib, dser, bits, flags;
  DBuf<double> drb;
  DBuf<unsigned long long> err;
  if (bits.alloc(2 * (size_t)L.nb) || flags.alloc(4) || err.alloc(1)) return FUSS_EHIP;
  HIPCHK(hipMemsetAsync(err.p, 0xff, sizeof(unsigned long long), c->stream));
  if (nf) {
    k_fill_i32<<<(unsigned)std::min<size_t>((nf + 255) / 256, 4096), 256, 0, c->stream>>>(L.fmap.p, (long)nf, -1);
    HIPCHK(hipGetLastError());
  }
  if (n) {
    if (dib.alloc(15 * (size_t)n) || drb.alloc(6 * (size_t)n) || (!L.hser.empty() && dser.alloc(n)))
      return FUSS_EHIP;
    e = upload(dib.p, L.hib.data(), sizeof(int) * 15 * (size_t)n);
    e |= upload(drb.p, L.hrb.data(), sizeof(double) * 6 * (size_t)n);
    if (!L.hser.empty()) e |= upload(dser.p, L.hser.data(), sizeof(int) * (size_t)n);
    if (e) return e;
    BcBuildArgs a{};
    a.ib = dib.p; a.rb = drb.p; a.ser = L.hser.empty() ? nullptr : dser.p;
    a.n = n; a.nb = L.nb; a.blk = L.blk.p; a.nodes = L.nodes.p;
    a.rec = L.bc.p; a.gm = L.gmap.p; a.fmap = L.fmap.p; a.bbits = bits.p; a.lflags = flags.p; a.err = err.p;
    k_build_bc<<<(n + 127) / 128, 128, 0, c->stream>>>(a);
    HIPCHK(hipGetLastError());
  }
  if (nf) {
    k_fmap_check<<<(unsigned)std::min<size_t>((nf + 255) / 256, 4096), 256, 0, c->stream>>>(L.fmap.p, (long)nf, n,
                                                                                           err.p);
    HIPCHK(hipGetLastError());
  }
  unsigned long long herr = 0;
  std::vector<int> hbits(2 * (size_t)L.nb);
  int hflags[4];
  HIPCHK(hipMemcpyAsync(&herr, err.p, sizeof(herr), hipMemcpyDeviceToHost, c->stream));
  HIPCHK(hipMemcpyAsync(hbits.data(), bits.p, sizeof(int) * hbits.size(), hipMemcpyDeviceToHost, c->stream));
  HIPCHK(hipMemcpyAsync(hflags, flags.p, sizeof(hflags), hipMemcpyDeviceToHost, c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  if (herr != ~0ULL) return (int)(herr & 0xff);
  L.any_active = 0;
  for (int b = 0; b < L.nb; b++) {
    L.hblk[b].face_active |= hbits[2 * b];
    L.hblk[b].conn |= hbits[2 * b + 1];
    L.any_active |= L.hblk[b].face_active;
  }
  L.no_xcon = !hflags[0];
  L.no_ycon = !hflags[1];
  L.no_zcon = !hflags[2];
  L.ghost_read = hflags[3];
  return 0;
}

int sync_status(fuss_ctx* c, int* bijk = nullptr) {
  HIPCHK(hipMemcpyAsync(c->hstatus, c->status.p, 8 * sizeof(int), hipMemcpyDeviceToHost, c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  unsigned long long key;
  memcpy(&key, c->hstatus + 6, sizeof(key));
  if (bijk) {
    if (c->hstatus[0] == FUSS_ESTATE && key != ~0ULL) {
      // first failing cell in the reference's loop order (flag_state)
      bijk[0] = (int)((key >> 48) & 0x1fff);
      bijk[1] = (int)(key & 0xffff);
      bijk[2] = (int)((key >> 16) & 0xffff);
      bijk[3] = (int)((key >> 32) & 0xffff);
    } else {
      for (int q = 0; q < 4; q++) bijk[q] = c->hstatus[1 + q];
    }
  }
  return c->hstatus[0];
}

// Host mirrors of a level's device clock (synchronises the stream).
int read_clock(fuss_ctx* c, Level& L) {
  HIPCHK(hipMemcpyAsync(c->hscal + 6, L.clk.p, sizeof(double), hipMemcpyDeviceToHost, c->stream));
  HIPCHK(hipMemcpyAsync(c->hcnt, L.cnt.p, 4 * sizeof(int), hipMemcpyDeviceToHost, c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  L.time = c->hscal[6];
  L.iter = c->hcnt[0];
  return 0;
}

int set_clock(fuss_ctx* c, Level& L, double time, int iter) {
  k_set_clock<<<1, 64, 0, c->stream>>>(L.clk.p, L.cnt.p, time, iter);
  HIPCHK(hipGetLastError());
  L.time = time;
  L.iter = iter;
  return 0;
}

}  // namespace

namespace fuss {

// Event pair around one launch of a kernel class (profiling mode only).
int prof_begin(fuss_ctx* c, int cls) {
  auto& P = c->prof;
  if (!P.enabled) return -1;
  if (2 * (P.used + 1) > P.pool.size()) {
    for (int q = 0; q < 256; q++) {
      hipEvent_t e;
      if (hipEventCreate(&e) != hipSuccess) return -1;
      P.pool.push_back(e);
    }
  }
  const int id = (int)P.used++;
  P.cls.resize(P.used);
  P.cls[id] = cls;
  (void)hipEventRecord(P.pool[2 * id], c->stream);
  return id;
}
void prof_end(fuss_ctx* c, int id) {
  if (id >= 0) (void)hipEventRecord(c->prof.pool[2 * id + 1], c->stream);
}

}  // namespace fuss

namespace {


template <typename R>
Tab<R> dtab(fuss_ctx* c);
template <>
Tab<double> dtab<double>(fuss_ctx* c) {
  Tab<double> t;
  t.cp = c->dcp.p; t.rho = c->drho.p; t.kap = c->dkap.p; t.h = c->dh.p;
  t.kap2 = c->dkap2.p; t.pr4 = c->dpr4.p;
  t.nT = c->nT; t.Ti1 = c->Ti1; t.Ti2 = c->Ti1 + c->nT - 1;
  return t;
}
// the FP32 copies of the tables (converted by f32_prepare)
template <>
Tab<float> dtab<float>(fuss_ctx* c) {
  Tab<float> t;
  t.cp = c->fcp.p; t.rho = c->frho.p; t.kap = c->fkap.p; t.h = c->fh.p;
  t.kap2 = c->fkap2.p; t.pr4 = c->fpr4.p;
  t.nT = c->nT; t.Ti1 = c->Ti1; t.Ti2 = c->Ti1 + c->nT - 1;
  return t;
}

template <typename R>
FieldView<R> fields(fuss_ctx* c, Level& L);
template <>
FieldView<double> fields<double>(fuss_ctx* c, Level& L) {
  FieldView<double> v;
  for (int q = 0; q < 3; q++) v.Tb[q] = L.Tb[q].p;
  v.res = L.R.p; v.res2 = L.R2.p; v.dt = L.dt.p; v.qvol = L.qvol.p; v.vol = L.vol.p; v.invL2 = L.invL2.p;
  for (int w = 0; w < 9; w++) v.W[w] = L.W[w].p;
  v.bcv = L.bcv.p; v.bcf = L.bcf.p;
  (void)c;
  return v;
}
template <>
FieldView<float> fields<float>(fuss_ctx* c, Level& L) {
  FieldView<float> v;
  for (int q = 0; q < 3; q++) v.Tb[q] = L.f.Tb[q].p;
  v.res = L.f.R.p; v.res2 = L.f.R2.p; v.dt = L.f.dt.p; v.qvol = L.f.qvol.p; v.vol = L.f.vol.p;
  v.invL2 = L.invL2.p;
  for (int w = 0; w < 9; w++) v.W[w] = L.f.W[w].p;
  v.bcv = L.f.bcv.p; v.bcf = L.f.bcf.p;
  (void)c;
  return v;
}

int convert_d2f(fuss_ctx* c, const double* src, float* dst, size_t n) {
  if (!n) return 0;
  k_d2f<<<(unsigned)std::min<size_t>((n + 255) / 256, 8192), 256, 0, c->stream>>>(src, dst, (long)n);
  return hipGetLastError() == hipSuccess ? 0 : FUSS_EHIP;
}
int convert_f2d(fuss_ctx* c, const float* src, double* dst, size_t n) {
  if (!n) return 0;
  k_f2d<<<(unsigned)std::min<size_t>((n + 255) / 256, 8192), 256, 0, c->stream>>>(src, dst, (long)n);
  return hipGetLastError() == hipSuccess ? 0 : FUSS_EHIP;
}

// FP32 mirrors of a level's static fields and of the tables (allocated and
// converted on the first FP32 step, the tables again after a change).
// Outside stream capture.
int f32_prepare(fuss_ctx* c, Level& L) {
  if (!c->ftab_ok) {
    const size_t n = (size_t)c->nmat * c->nT;
    if (c->fcp.n != n && (c->fcp.alloc(n) || c->frho.alloc(n) || c->fkap.alloc(n) || c->fh.alloc(n) ||
                          c->fkap2.alloc(2 * n) || c->fpr4.alloc(4 * n)))
      return FUSS_EHIP;
    int e = convert_d2f(c, c->dcp.p, c->fcp.p, n) | convert_d2f(c, c->drho.p, c->frho.p, n) |
            convert_d2f(c, c->dkap.p, c->fkap.p, n) | convert_d2f(c, c->dh.p, c->fh.p, n) |
            convert_d2f(c, c->dkap2.p, c->fkap2.p, 2 * n) | convert_d2f(c, c->dpr4.p, c->fpr4.p, 4 * n);
    if (e) return FUSS_EHIP;
    c->ftab_ok = 1;
  }
  if (L.f_ready) return 0;
  const size_t n = (size_t)L.npad;
  auto& F = L.f;
  for (auto& t : F.Tb)
    if (t.alloc(n)) return FUSS_EHIP;
  if (F.R.alloc(n) || F.R2.alloc(n) || F.dt.alloc(n) || F.qvol.alloc(n) || F.vol.alloc(n)) return FUSS_EHIP;
  for (int w = 0; w < 9; w++)
    if (F.W[w].alloc(L.W[w].n)) return FUSS_EHIP;
  if (F.bcv.alloc(L.bcv.n) || F.bcf.alloc(L.bcf.n)) return FUSS_EHIP;
  int e = convert_d2f(c, L.qvol.p, F.qvol.p, n) | convert_d2f(c, L.vol.p, F.vol.p, n);
  for (int w = 0; w < 9; w++) e |= convert_d2f(c, L.W[w].p, F.W[w].p, L.W[w].n);
  if (e) return FUSS_EHIP;
  L.f_ready = 1;
  return 0;
}

// The level's current state into the FP32 buffers (ghosts included), or
// back into the FP64 ones; the next step then computes its dt afresh.
int state_to_f32(fuss_ctx* c, Level& L) {
  int e = f32_prepare(c, L);
  if (e || L.state_f) return e;
  e = convert_d2f(c, L.Tb[L.cur].p, L.f.Tb[L.cur].p, L.npad);
  L.state_f = 1;
  L.dt_next = 0;
  return e;
}
int state_to_f64(fuss_ctx* c, Level& L) {
  if (!L.state_f) return 0;
  const int e = convert_f2d(c, L.f.Tb[L.cur].p, L.Tb[L.cur].p, L.npad);
  L.state_f = 0;
  L.dt_next = 0;
  return e;
}
int all_to_f64(fuss_ctx* c) {
  for (auto& L : c->lev) {
    const int e = state_to_f64(c, L);
    if (e) return e;
  }
  return 0;
}

template <typename R>
int launch_ghost(fuss_ctx* c, Level& L, R* T) {
  if (!L.nbound) return 0;
  const int pid = prof_begin(c, FUSS_K_GHOST);
  k_ghost<R><<<(L.nbound + 255) / 256, 256, 0, c->stream>>>(T, L.gmap.p, L.nbound, L.cnt.p + 1);
  if constexpr (std::is_same<R, double>::value) {
    if (L.nchim)
      k_chimera<<<(L.nchim + 127) / 128, 128, 0, c->stream>>>(T, L.chr.p, L.chd.p, L.chv.p, L.nchim, L.mat.p,
                                                              dtab<double>(c), c->status.p, L.cnt.p + 1);
  }
  prof_end(c, pid);
  return hipGetLastError() == hipSuccess ? 0 : FUSS_EHIP;
}

int restrict_level(fuss_ctx* c, int fine) {
  Level& F = c->lev[fine];
  Level& C = c->lev[fine + 1];
  for (int b = 0; b < F.nb; b++) {
    MGArgs a{};
    a.F = F.hblk[b]; a.Cb = C.hblk[b];
    a.fT = F.Tb[F.cur].p; a.fmat = F.mat.p; a.fvol = F.vol.p; a.fqvol = F.qvol.p;
    a.cT = C.Tb[C.cur].p; a.cmat = C.mat.p; a.cvol = C.vol.p; a.cqvol = C.qvol.p;
    a.tab = dtab<double>(c); a.status = c->status.p; a.bidx = b;
    const long n = (long)a.Cb.ni * a.Cb.nj * a.Cb.nk;
    k_restrict<<<(unsigned)((n + 255) / 256), 256, 0, c->stream>>>(a);
    if (hipGetLastError() != hipSuccess) return FUSS_EHIP;
  }
  return 0;
}

}  // namespace

// Staging of fuss_wall_noalloc: Tw and qw of up to FUSS_WALL_NOALLOC_MAX
// entries (4 MiB on every GPU the library runs on, held until the process
// ends).  Host code uses the per-device address from hipGetSymbolAddress,
// never the array's name.
static __device__ __align__(256) double scratch[2 * FUSS_WALL_NOALLOC_MAX];

namespace {
constexpr int MAX_DEVICES = 64;
std::mutex g_scratch_mu;
double* g_scratch[MAX_DEVICES] = {};

// Device address of `scratch` on the current device, looked up once per
// device (fuss_create: the library's set-up call).
int scratch_setup(int dev) {
  if (dev < 0 || dev >= MAX_DEVICES) return FUSS_EHIP;
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (g_scratch[dev]) return 0;
  void* ptr = nullptr;
  if (hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)) != hipSuccess || !ptr) return FUSS_EHIP;
  g_scratch[dev] = static_cast<double*>(ptr);
  return 0;
}
}  // namespace

// ------------------------------------------------------------------------
// C-ABI
// ------------------------------------------------------------------------
extern "C" {

const char* fuss_strerror(int code) {
  switch (code) {
    case FUSS_OK: return "ok";
    case FUSS_EARG: return "invalid argument or call order";
    case FUSS_ESTATE: return "NaN or T<0 detected";
    case FUSS_ETABLE: return "temperature or enthalpy outside the property table";
    case FUSS_EMETRIC: return "metric tensor det=0 or negative volume";
    case FUSS_EUNSUPPORTED: return "unsupported feature for this case";
    case FUSS_EHIP: return "HIP runtime error";
    case FUSS_EBC: return "boundary list inconsistent with the blocks";
  }
  return "unknown status";
}

int fuss_create(int nlevels, int nblocks, int precision, fuss_ctx** out) {
  if (!out || nlevels < 1 || nblocks < 1 || (precision != FUSS_FP64 && precision != FUSS_FP32)) return FUSS_EARG;
  int ndev = 0, dev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) return FUSS_EHIP;
  if (hipGetDevice(&dev) != hipSuccess || scratch_setup(dev)) return FUSS_EHIP;
  fuss_ctx* c = new fuss_ctx();
  c->dev = dev;
  c->nlev = nlevels;
  c->nb = nblocks;
  c->precision = precision;
  c->lev.resize(nlevels);
  for (auto& L : c->lev) {
    L.nb = nblocks;
    L.hb.resize(nblocks);
  }
  if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess) {
    delete c;
    return FUSS_EHIP;
  }
  c->own_stream = true;
  if (c->status.alloc(8) || c->scal.alloc(4) || c->run_hist.alloc(2 * RUN_G) ||
      hipHostMalloc((void**)&c->hscal, 8 * sizeof(double), 0) != hipSuccess ||
      hipHostMalloc((void**)&c->hstatus, 8 * sizeof(int), 0) != hipSuccess ||
      hipHostMalloc((void**)&c->hrun, 2 * RUN_G * sizeof(double), 0) != hipSuccess ||
      hipHostMalloc((void**)&c->hcnt, 8 * sizeof(int), 0) != hipSuccess) {
    fuss_destroy(c);
    return FUSS_EHIP;
  }
  // orthogonal 3-D stages: the LDS-ring kernel (default) or, with
  // FUSS_STAGE_ROW=1, the barrier-free row march
  const char* ev = getenv("FUSS_STAGE_ROW");
  c->stage_lds = !(ev && ev[0] == '1');
  *out = c;
  return 0;
}

int fuss_set_table(fuss_ctx* c, int nmat, int Ti1, int nT, const double* cp, const double* rho,
                   const double* kappa, const double* h) {
  if (c) drop_dt_next(c);
  // rows start at a non-negative temperature (Load_Table reads rows from
  // T = 1 K): the kernels' row test int(T) < Ti2 then equals T < Ti2
  if (!c || nmat < 1 || nT < 2 || Ti1 < 0 || !cp || !rho || !kappa || !h) return FUSS_EARG;
  const size_t n = (size_t)nmat * nT;
  c->nmat = nmat; c->Ti1 = Ti1; c->nT = nT;
  c->tcp.assign(cp, cp + n); c->trho.assign(rho, rho + n); c->tkap.assign(kappa, kappa + n); c->th.assign(h, h + n);
  for (int m = 0; m < nmat; m++)
    for (int t = 1; t < nT; t++)
      if (h[(size_t)m * nT + t] < h[(size_t)m * nT + t - 1]) return FUSS_ETABLE;
  int e = c->dcp.alloc(n) | c->drho.alloc(n) | c->dkap.alloc(n) | c->dh.alloc(n);
  if (e) return FUSS_EHIP;
  e |= upload(c->dcp.p, cp, n * 8);
  e |= upload(c->drho.p, rho, n * 8);
  e |= upload(c->dkap.p, kappa, n * 8);
  e |= upload(c->dh.p, h, n * 8);
  // (v, v(T+1)-v(T)) rows; the last row has no successor and is never used
  std::vector<double> k2(2 * n, 0.0), p4(4 * n, 0.0);
  for (int m = 0; m < nmat; m++)
    for (int t = 0; t < nT; t++) {
      const size_t q = (size_t)m * nT + t;
      const bool nx = t + 1 < nT;
      k2[2 * q] = kappa[q];
      k2[2 * q + 1] = nx ? kappa[q + 1] - kappa[q] : 0.0;
      p4[4 * q] = rho[q];
      p4[4 * q + 1] = nx ? rho[q + 1] - rho[q] : 0.0;
      p4[4 * q + 2] = cp[q];
      p4[4 * q + 3] = nx ? cp[q + 1] - cp[q] : 0.0;
    }
  if (c->dkap2.alloc(2 * n) || c->dpr4.alloc(4 * n)) return FUSS_EHIP;
  e |= upload(c->dkap2.p, k2.data(), k2.size() * 8);
  e |= upload(c->dpr4.p, p4.data(), p4.size() * 8);
  return e;
}

int fuss_set_block(fuss_ctx* c, int b, int ni, int nj, int nk, const double* nodes, const double* T,
                   const double* matID, const double* qvol) {
  if (!c || b < 1 || b > c->nb || ni < 1 || nj < 1 || nk < 1 || !nodes || !T || !matID) return FUSS_EARG;
  HostBlock& h = c->lev[0].hb[b - 1];
  h.n[0] = ni; h.n[1] = nj; h.n[2] = nk;
  const size_t nn = 3 * (size_t)(ni + 1) * (nj + 1) * (nk + 1), nc = (size_t)ni * nj * nk;
  h.nodes.assign(nodes, nodes + nn);
  h.T.assign(T, T + nc);
  h.matID.assign(matID, matID + nc);
  if (qvol) h.qvol.assign(qvol, qvol + nc);
  else h.qvol.assign(nc, 0.0);
  return 0;
}

int fuss_add_series(fuss_ctx* c, int periodic, int n, const double* t, const double* v, int* id) {
  if (!c || n < 1 || !t || !v || (int)c->series.size() >= MAX_SERIES) return FUSS_EARG;
  fuss_ctx::Series s;
  s.periodic = periodic;
  s.t.assign(t, t + n);
  s.v.assign(v, v + n);
  c->series.push_back(s);
  if (id) *id = (int)c->series.size() - 1;
  return 0;
}

int fuss_set_bc(fuss_ctx* c, int level, int nbound, const int* ib, const double* rb, const int* series) {
  if (!c || level < 1 || level > c->nlev || nbound < 0 || (nbound && (!ib || !rb))) return FUSS_EARG;
  Level& L = c->lev[level - 1];
  L.nbound = nbound;
  L.hib.assign(ib, ib + 15 * (size_t)nbound);
  L.has_conn = 0;
  for (int e = 0; e < nbound; e++) L.has_conn |= ib[15 * (size_t)e + 5] == 101 || ib[15 * (size_t)e + 5] == 102;
  L.hrb.assign(rb, rb + 6 * (size_t)nbound);
  if (series) L.hser.assign(series, series + nbound);
  else L.hser.clear();
  for (int e = 0; e < nbound && series; e++)
    if (series[e] >= (int)c->series.size()) return FUSS_EARG;
  return 0;
}

int fuss_set_chimera(fuss_ctx* c, int level, int n, const int* entry, const int* n1, const int* n2,
                     const int* donors, const double* vf) {
  if (!c || c->setup_done || level < 1 || level > c->nlev || n < 0 || (n && (!entry || !n1 || !n2))) return FUSS_EARG;
  Level& L = c->lev[level - 1];
  size_t m = 0;
  for (int q = 0; q < n; q++) {
    if (entry[q] < 0 || entry[q] >= L.nbound || L.hib[15 * (size_t)entry[q] + 5] != 102 || n1[q] < 1 || n2[q] < 0)
      return FUSS_EARG;
    m += (size_t)(n1[q] + n2[q]);
  }
  if (m && (!donors || !vf)) return FUSS_EARG;
  L.hch_ent.assign(entry, entry + n);
  L.hch_n1.assign(n1, n1 + n);
  L.hch_n2.assign(n2, n2 + n);
  L.hch_don.assign(donors, donors + 4 * m);
  L.hch_vf.assign(vf, vf + m);
  return 0;
}

int fuss_setup(fuss_ctx* c, double initial_time) {
  if (!c || c->setup_done || c->tkap.empty()) return FUSS_EARG;
  for (int b = 0; b < c->nb; b++)
    if (c->lev[0].hb[b].nodes.empty()) return FUSS_EARG;
  // Check_Mesh_Type, Lib_Metrics.f90:13-66
  const HostBlock& b1 = c->lev[0].hb[0];
  c->ndir = b1.n[2] > 1 ? 3 : (b1.n[1] == 1 ? 1 : 2);
  // Check_Multigrid + Setup_Multigrid/Coarse_Grid (Lib_Multigrid.f90:11-35, :223-265)
  const int rap = 1 << (c->nlev - 1);
  for (int b = 0; b < c->nb; b++) {
    const HostBlock& h = c->lev[0].hb[b];
    int chk = 0;
    for (int d = 0; d < 3; d++) {
      if (h.n[d] % rap == 0) chk++;
      if (d == 2 && h.n[d] == 1) chk++;
    }
    if (chk < 3) return FUSS_EARG;
  }
  for (int m = 1; m < c->nlev; m++)
    for (int b = 0; b < c->nb; b++) {
      const HostBlock& F = c->lev[m - 1].hb[b];
      HostBlock& C = c->lev[m].hb[b];
      C.n[0] = F.n[0] / 2; C.n[1] = F.n[1] / 2; C.n[2] = std::max(1, F.n[2] / 2);
      C.nodes.assign(3 * (size_t)(C.n[0] + 1) * (C.n[1] + 1) * (C.n[2] + 1), 0.0);
      const int step = 2 - F.n[2] % 2;
      for (int k = 0; k <= F.n[2]; k += step)
        for (int j = 0; j <= F.n[1]; j += 2)
          for (int i = 0; i <= F.n[0]; i += 2) {
            const int k2 = F.n[2] == 1 ? k : k / 2;
            const double* s = &F.nodes[3 * (i + (size_t)(F.n[0] + 1) * (j + (size_t)(F.n[1] + 1) * k))];
            double* d = &C.nodes[3 * (i / 2 + (size_t)(C.n[0] + 1) * (j / 2 + (size_t)(C.n[1] + 1) * k2))];
            d[0] = s[0]; d[1] = s[1]; d[2] = s[2];
          }
    }
  HIPCHK(hipMemsetAsync(c->status.p, 0, 6 * sizeof(int), c->stream));
  HIPCHK(hipMemsetAsync(c->status.p + 6, 0xff, 2 * sizeof(int), c->stream));  // no failing cell yet
  for (int m = 0; m < c->nlev; m++) {
    Level& L = c->lev[m];
    // 2-D levels keep one plane on each side: it stores the k-face data
    L.kp = c->ndir == 2 ? 1 : GC;
    int e = build_level_layout(c, L);
    if (e) return e;
    // BC records: on the device, or on the host for a level with chimera
    // entries (donor lists, extrapolated ghost metrics); both fill the face
    // map offsets and the face activity into hblk
    const bool host_bc = level_has_chimera(L);
    if (host_bc) {
      e = build_bc(c, L);
      if (e) return e;
    }
    e = alloc_level(c, L);
    if (e) return e;
    e = L.blk.alloc(L.nb);
    if (e) return e;
    e = upload(L.tiles.p, L.htiles.data(), sizeof(Tile) * L.htiles.size());
    if (!L.hctiles.empty()) e |= upload(L.ctiles.p, L.hctiles.data(), sizeof(Tile) * L.hctiles.size());
    if (!L.hc8tiles.empty()) e |= upload(L.c8tiles.p, L.hc8tiles.data(), sizeof(Tile) * L.hc8tiles.size());
    e |= upload(L.xytiles.p, L.hxytiles.data(), sizeof(Tile) * L.hxytiles.size());
    // nodes, each block's straight into its place in the arena
    for (int b = 0; b < L.nb && !e; b++)
      e |= upload(L.nodes.p + 3 * L.hblk[b].node_off, L.hb[b].nodes.data(), L.hb[b].nodes.size() * 8);
    if (!host_bc) {
      const size_t nf = face_map_offsets(L);
      e |= upload(L.blk.p, L.hblk.data(), sizeof(BlockInfo) * L.nb);
      if (e) return e;
      e = build_bc_dev(c, L, nf);
      if (e) return e;
    }
    e |= upload(L.blk.p, L.hblk.data(), sizeof(BlockInfo) * L.nb);
    if (e) return e;
    // cell metrics
    GeomArgs g{};
    g.blk = L.blk.p; g.tiles = L.tiles.p; g.nodes = L.nodes.p; g.M = L.M.p; g.vol = L.vol.p;
    g.invL2 = L.invL2.p; g.ndir = c->ndir; g.status = c->status.p;
    k_cell_metrics<<<(unsigned)L.htiles.size(), TILE_T, 0, c->stream>>>(g);
    HIPCHK(hipGetLastError());
    // faces + weights per block and direction
    DBuf<int> tang;
    if (tang.alloc(L.nb)) return FUSS_EHIP;
    for (int b = 0; b < L.nb; b++)
      for (int D = 0; D < 3; D++) {
        FaceArgs f{};
        f.B = L.hblk[b]; f.bidx = b; f.D = D; f.sign = norm_sign(L.hb[b], D);
        f.nodes = L.nodes.p; f.M = L.M.p; f.F = L.F[D].p;
        for (int cc = 0; cc < 3; cc++) f.W[cc] = L.W[3 * D + cc].p;
        f.tang = tang.p + b;
        const long nf = (long)(f.B.ni + (D == 0)) * (f.B.nj + (D == 1)) * (f.B.nk + (D == 2));
        k_face_metrics<<<(unsigned)((nf + 255) / 256), 256, 0, c->stream>>>(f);
        HIPCHK(hipGetLastError());
      }
    std::vector<int> htang(L.nb);
    HIPCHK(hipMemcpyAsync(htang.data(), tang.p, sizeof(int) * L.nb, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    tang.free();
    L.tang = 0;
    for (int b = 0; b < L.nb; b++) {
      L.hblk[b].tang = htang[b];
      L.tang |= htang[b];
    }
    HIPCHK(hipMemcpy(L.blk.p, L.hblk.data(), sizeof(BlockInfo) * L.nb, hipMemcpyHostToDevice));
    e = rect_prepare(c, L);
    if (e) return e;
    if (L.nbound) {
      k_bc_geometry<<<(L.nbound + 255) / 256, 256, 0, c->stream>>>(L.bc.p, L.nbound, L.blk.p, L.M.p, L.F[0].p,
                                                                  L.F[1].p, L.F[2].p);
      HIPCHK(hipGetLastError());
    }
    L.time = initial_time;
    L.iter = 0;
    k_set_clock<<<1, 64, 0, c->stream>>>(L.clk.p, L.cnt.p, initial_time, 0);
    HIPCHK(hipGetLastError());
  }
  // time series of 301/302 values, evaluated on the device at the level time
  {
    std::vector<double> st, sv;
    std::vector<int> so, sn, sp;
    for (const auto& q : c->series) {
      so.push_back((int)st.size());
      sn.push_back((int)q.t.size());
      sp.push_back(q.periodic);
      st.insert(st.end(), q.t.begin(), q.t.end());
      sv.insert(sv.end(), q.v.begin(), q.v.end());
    }
    const size_t ns = std::max<size_t>(so.size(), 1), np_ = std::max<size_t>(st.size(), 1);
    if (c->ser_t.alloc(np_) || c->ser_v.alloc(np_) || c->ser_off.alloc(ns) || c->ser_n.alloc(ns) ||
        c->ser_per.alloc(ns))
      return FUSS_EHIP;
    if (!st.empty()) {
      HIPCHK(hipMemcpy(c->ser_t.p, st.data(), st.size() * 8, hipMemcpyHostToDevice));
      HIPCHK(hipMemcpy(c->ser_v.p, sv.data(), sv.size() * 8, hipMemcpyHostToDevice));
      HIPCHK(hipMemcpy(c->ser_off.p, so.data(), so.size() * 4, hipMemcpyHostToDevice));
      HIPCHK(hipMemcpy(c->ser_n.p, sn.data(), sn.size() * 4, hipMemcpyHostToDevice));
      HIPCHK(hipMemcpy(c->ser_per.p, sp.data(), sp.size() * 4, hipMemcpyHostToDevice));
    }
  }
  // level-1 fields: T, material index, qvol, scattered on the device from
  // each block's arrays
  {
    Level& L = c->lev[0];
    size_t ncmax = 1;
    for (int b = 0; b < L.nb; b++) ncmax = std::max(ncmax, L.hb[b].T.size());
    DBuf<double> sT, sq, sm;
    DBuf<int> mm, ferr;
    if (sT.alloc(ncmax) || sq.alloc(ncmax) || sm.alloc(ncmax) || mm.alloc(2 * (size_t)L.nb) || ferr.alloc(1))
      return FUSS_EHIP;
    std::vector<int> hmm(2 * (size_t)L.nb);
    for (int b = 0; b < L.nb; b++) {
      hmm[2 * b] = 0x7fffffff;
      hmm[2 * b + 1] = -0x7fffffff;
    }
    int e = upload(mm.p, hmm.data(), sizeof(int) * hmm.size());
    for (int b = 0; b < L.nb && !e; b++) {
      const HostBlock& h = L.hb[b];
      const size_t nc = h.T.size();
      // the staging buffers are reused block after block: each upload waits
      // for the previous block's scatter
      HIPCHK(hipStreamSynchronize(c->stream));
      e |= upload(sT.p, h.T.data(), nc * 8);
      e |= upload(sq.p, h.qvol.data(), nc * 8);
      e |= upload(sm.p, h.matID.data(), nc * 8);
      if (e) break;
      k_scatter_fields<<<(unsigned)std::min<size_t>((nc + 255) / 256, 8192), 256, 0, c->stream>>>(
          L.hblk[b], sT.p, sq.p, sm.p, L.Tb[L.cur].p, L.qvol.p, L.mat.p, c->nmat, mm.p + 2 * b, ferr.p);
      HIPCHK(hipGetLastError());
    }
    if (e) return e;
    int herr = 0;
    HIPCHK(hipMemcpyAsync(hmm.data(), mm.p, sizeof(int) * hmm.size(), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipMemcpyAsync(&herr, ferr.p, sizeof(int), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    if (herr) return FUSS_EARG;
    for (int b = 0; b < L.nb; b++) {
      // a block of one material keeps it on every level (restriction copies
      // a child's material, Lib_Multigrid.f90:113)
      const int umat = hmm[2 * b] == hmm[2 * b + 1] ? hmm[2 * b] : -1;
      for (int lv = 0; lv < c->nlev; lv++) c->lev[lv].hblk[b].umat = umat;
    }
    for (int lv = 0; lv < c->nlev; lv++) {
      Level& M = c->lev[lv];
      M.all_umat = 1;
      for (int b = 0; b < M.nb; b++) M.all_umat &= M.hblk[b].umat >= 0;
    }
    for (int lv = 0; lv < c->nlev; lv++)
      HIPCHK(hipMemcpy(c->lev[lv].blk.p, c->lev[lv].hblk.data(), sizeof(BlockInfo) * c->lev[lv].nb,
                       hipMemcpyHostToDevice));
  }
  // matIDg of connections: material of the donor cell (Fill_matIDg after the
  // first Ghost_Connection copy, Lib_Ghost.f90:149-153, :436-470); chimera
  // ghosts got theirs from the donor list
  {
    Level& L = c->lev[0];
    if (L.nbound) {
      k_set_matg<<<(L.nbound + 255) / 256, 256, 0, c->stream>>>(L.bc.p, L.nbound, L.mat.p);
      HIPCHK(hipGetLastError());
    }
  }
  int e = launch_ghost(c, c->lev[0], c->lev[0].Tb[c->lev[0].cur].p);
  if (e) return e;
  for (int m = 1; m < c->nlev; m++) {
    e = restrict_level(c, m - 1);
    if (e) return e;
    // coarse matIDg: donor material from the restricted material field
    Level& L = c->lev[m];
    if (L.nbound) {
      k_set_matg<<<(L.nbound + 255) / 256, 256, 0, c->stream>>>(L.bc.p, L.nbound, L.mat.p);
      HIPCHK(hipGetLastError());
    }
    e = launch_ghost(c, L, L.Tb[L.cur].p);
    if (e) return e;
  }
  // blocks of uniform qvol (after restriction, on every level)
  for (int m = 0; m < c->nlev; m++) {
    Level& L = c->lev[m];
    DBuf<int> nonu;
    if (nonu.alloc(L.nb)) return FUSS_EHIP;
    k_qvol_nonuniform<<<dim3(L.nb, 64), 256, 0, c->stream>>>(L.blk.p, L.qvol.p, nonu.p);
    HIPCHK(hipGetLastError());
    std::vector<int> hn(L.nb);
    HIPCHK(hipMemcpyAsync(hn.data(), nonu.p, sizeof(int) * L.nb, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    L.all_uqv = 1;
    for (int b = 0; b < L.nb; b++) {
      L.hblk[b].uqv = !hn[b];
      L.all_uqv &= L.hblk[b].uqv;
    }
    HIPCHK(hipMemcpy(L.blk.p, L.hblk.data(), sizeof(BlockInfo) * L.nb, hipMemcpyHostToDevice));
  }
  // compact wall records, now that every level's materials are set
  for (int m = 0; m < c->nlev; m++) {
    Level& L = c->lev[m];
    if (!L.nbound) continue;
    k_bc_hot<<<(L.nbound + 255) / 256, 256, 0, c->stream>>>(L.bc.p, L.nbound, L.mat.p, L.bch.p);
    HIPCHK(hipGetLastError());
    // compact records of the connection entries (their geometry and ghost
    // materials are final now)
    std::vector<int> ent;
    for (int e = 0; e < L.nbound; e++)
      if (L.hib[15 * (size_t)e + 5] == 101) ent.push_back(e);
    L.nconn = (int)ent.size();
    if (L.nconn) {
      DBuf<int> dent;
      if (dent.alloc(ent.size()) || L.cd.alloc((size_t)CONN_ND * L.nconn) || L.ci.alloc((size_t)CONN_NI * L.nconn))
        return FUSS_EHIP;
      if (upload(dent.p, ent.data(), sizeof(int) * ent.size())) return FUSS_EHIP;
      k_conn_build<<<(L.nconn + 255) / 256, 256, 0, c->stream>>>(L.bc.p, dent.p, L.nconn, L.blk.p, L.cd.p, L.ci.p);
      HIPCHK(hipGetLastError());
      HIPCHK(hipStreamSynchronize(c->stream));
    }
  }
  int bijk[4];
  const int st = sync_status(c, bijk);
  if (st) return st;
  c->setup_done = 1;
  return 0;
}

}  // extern "C"

namespace {

// (stage and IRS launchers: stage_*.hip, irs_*.hip)


// Line tiles and elimination factors of the tridiagonal IRS for beta (host
// arithmetic in the oracle's order: den_1 = 1+2b, den_s = (1+2b) + b c_{s-1},
// c_s = -b / den_s).  Synchronous uploads: call outside stream capture.
int tri_prepare(fuss_ctx* c, Level& L, double beta) {
  if (L.tri_beta == beta && L.tri.p) return 0;
  if (L.htx.empty()) {
    for (int b = 0; b < L.nb; b++) {
      const BlockInfo& B = L.hblk[b];
      for (int k = 1; k <= B.nk; k++)
        for (int j = 1; j <= B.nj; j += TILE_X) L.htx.push_back(Tile{b, 1, j, k});
      for (int k = 1; k <= B.nk; k += TILE_Y)
        for (int i = 1; i <= B.ni; i += TILE_X) L.hty.push_back(Tile{b, i, 1, k});
      for (int j = 1; j <= B.nj; j += TILE_Y)
        for (int i = 1; i <= B.ni; i += TILE_X) L.htz.push_back(Tile{b, i, j, 1});
    }
    if (L.ttx.alloc(L.htx.size()) || L.tty.alloc(L.hty.size()) || L.ttz.alloc(L.htz.size())) return FUSS_EHIP;
    int e = upload(L.ttx.p, L.htx.data(), L.htx.size() * sizeof(Tile));
    e |= upload(L.tty.p, L.hty.data(), L.hty.size() * sizeof(Tile));
    e |= upload(L.ttz.p, L.htz.data(), L.htz.size() * sizeof(Tile));
    if (e) return e;
  }
  std::vector<double2> tab;
  std::vector<int> off(3 * L.nb);
  {
#pragma clang fp contract(off)
    const double bb = 1.0 + 2.0 * beta;
    for (int b = 0; b < L.nb; b++)
      for (int d = 0; d < 3; d++) {
        const BlockInfo& B = L.hblk[b];
        const int n = d == 0 ? B.ni : (d == 1 ? B.nj : B.nk);
        off[3 * b + d] = (int)tab.size();
        double cp = 0.0;
        for (int s = 1; s <= n; s++) {
          const double den = s == 1 ? bb : bb + beta * cp;
          cp = -beta / den;
          tab.push_back(make_double2(den, cp));
        }
      }
  }
  if (L.tri.alloc(std::max<size_t>(tab.size(), 1)) || L.trioff.alloc(off.size())) return FUSS_EHIP;
  int e = upload(L.tri.p, tab.data(), tab.size() * sizeof(double2));
  e |= upload(L.trioff.p, off.data(), off.size() * sizeof(int));
  if (e) return e;
  L.tri_beta = beta;
  (void)c;
  return 0;
}


const double RKC[3][3] = {{1.0, 0.0, 0.0}, {1.0, 0.50, 0.0}, {1.0, 0.25, 2.0 / 3.0}};

// 1/d when d is a (normal) power of two, so that x*(1/d) == x/d for every
// finite x away from the subnormal range; else 0 (divide).
double exact_inverse(double d) {
  int e = 0;
  const double m = frexp(d, &e);
  return (m == 0.5 && d > 1e-300 && d < 1e300) ? 1.0 / d : 0.0;
}
const double RK4A[4] = {0.25, 1.0 / 3.0, 0.5, 1.0};

// Explicit_Step on one level, enqueued on the stream (no host sync except
// the dtglobal readback of time-accurate runs with time-dependent BCs).
// The host side of External_Function (fuss_step_hook): the hook and its
// user pointer, called once per block and stage.
struct StageHook {
  fuss_stage_fn fn;
  void* user;
};

// Hands each block's raw R (and the stage input T) of the level to the hook
// in the block's i-fastest order and puts the returned R back.
int run_hook(fuss_ctx* c, Level& L, int level, int irk, const double* Tin, const StageHook* hk) {
  for (int b = 0; b < L.nb; b++) {
    const BlockInfo& B = L.hblk[b];
    const size_t n = (size_t)B.ni * B.nj * B.nk;
    std::vector<double> padR(B.npad), padT(B.npad), R(n), T(n);
    HIPCHK(hipMemcpyAsync(padR.data(), L.R.p + B.off, B.npad * 8, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipMemcpyAsync(padT.data(), Tin + B.off, B.npad * 8, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    for (int k = 1; k <= B.nk; k++)
      for (int j = 1; j <= B.nj; j++)
        for (int i = 1; i <= B.ni; i++) {
          const size_t q = (i - 1) + (size_t)B.ni * ((j - 1) + (size_t)B.nj * (k - 1));
          R[q] = padR[cidx(B, i, j, k) - B.off];
          T[q] = padT[cidx(B, i, j, k) - B.off];
        }
    hk->fn(hk->user, level, irk, b + 1, B.ni, B.nj, B.nk, T.data(), R.data());
    for (int k = 1; k <= B.nk; k++)
      for (int j = 1; j <= B.nj; j++)
        for (int i = 1; i <= B.ni; i++)
          padR[cidx(B, i, j, k) - B.off] = R[(i - 1) + (size_t)B.ni * ((j - 1) + (size_t)B.nj * (k - 1))];
    HIPCHK(hipMemcpyAsync(L.R.p + B.off, padR.data(), B.npad * 8, hipMemcpyHostToDevice, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
  }
  return 0;
}

template <typename R>
int enqueue_step_t(fuss_ctx* c, int level, const fuss_step_params* p, const StageHook* hk) {
  Level& L = c->lev[level - 1];
  const FieldView<R> V = fields<R>(c, L);
  const int nrk = p->n_rk;
  if (nrk < 1 || nrk > 4) return FUSS_EARG;
  const bool prim = p->integration == FUSS_PRIM;
  const bool irs = p->irs_beta > 0.0;
  if (irs && p->irs_mode != FUSS_IRS_JACOBI2 && p->irs_mode != FUSS_IRS_TRIDIAG) return FUSS_EARG;
  const bool tri = irs && p->irs_mode == FUSS_IRS_TRIDIAG;
  if (tri) {
    const int e = tri_prepare(c, L, p->irs_beta);  // no-op once prepared (fuss_run prepares before capture)
    if (e) return e;
  }
  const unsigned ntile = (unsigned)L.htiles.size();
  const int* halt = L.cnt.p + 1;
  // The fused x-y-z IRS pass of the last stage can leave the next step's
  // local dt (and its per-tile minima) behind: without a vnn ramp, when the
  // state is the one that pass wrote and vnn is unchanged.
  // A steady primitive step on a rectilinear 3-D level whose stages run the
  // LDS-ring kernel computes dt in its first stage instead (sdt): that
  // stage's input is T^n, and the cell's rho/cp row is loaded there anyway.
  const bool xyz = irs && !tri && L.no_ycon && L.no_zcon && c->ndir == 3 && L.kp == GC;
  const bool ring = L.kp == GC && !L.tang && L.dims2 && !(L.row_march >= 0 ? L.row_march == 1 : !c->stage_lds);
  // the stage kernel's columns (partials and dt minima per column)
  const unsigned nstcol = (unsigned)(ring && L.stage_ty == 8 ? L.hc8tiles.size() : L.hctiles.size());
  const int stlist = ring && L.stage_ty == 8 ? 3 : 1;
  const bool sdt = !hk && ring && L.rect_w && !c->rect_off && prim && !p->time_accurate && p->rampa_iter <= 1 &&
                   !c->adj_store_on;
  const bool fuse_dt = !hk && !sdt && xyz && p->rampa_iter <= 1 && !c->adj_store_on;
  const bool have_dt = fuse_dt && L.dt_next && L.dt_vnn == p->vnn;
  // latency-bound 2-D levels without IRS or connections: fused stages (BC
  // values in place, mirror ghosts written with each new state, the local
  // dt in the first stage of steady steps) and one kernel closing the step
  const bool fused = c->fuse_small && L.kp != GC && !irs && !hk && !c->adj_store_on && !L.has_conn && !L.nchim &&
                     ntile <= 16384;
  const bool fdt = fused && !p->time_accurate;
  unsigned nmin = ntile;
  int pid = -1;
  if (sdt) {
    nmin = nstcol;
  } else if (fdt) {
    // (the first stage computes it)
  } else if (have_dt) {
    // minima of the fused IRS pass's tiles (its list when dt was written)
    nmin = (unsigned)(L.dt_ty == 8 ? L.hc8tiles.size() : L.hctiles.size());
  } else {
    // Compute_dt + min (Mod_dt.f90:10-84); dtglobal starts at 1e5 (Mod_Explicit.f90:40)
    DtArgs<R> da{};
    da.blk = L.blk.p; da.tiles = L.tiles.p; da.T = V.Tb[L.cur]; da.mat = L.mat.p; da.invL2 = V.invL2;
    da.dt = V.dt; da.tab = dtab<R>(c); da.vnn = p->vnn; da.cnt = L.cnt.p; da.rampa = p->rampa_iter;
    da.pmin = L.pmin.p; da.status = c->status.p;
    pid = prof_begin(c, FUSS_K_DT);
    k_dt<R><<<ntile, TILE_T, 0, c->stream>>>(da);
    prof_end(c, pid);
    HIPCHK(hipGetLastError());
  }
  // device scalars: [0] dtglobal, [1] time, [2] Sum (T-TO)^2
  auto dt_final = [&]() -> int {
    pid = prof_begin(c, FUSS_K_REDUCE);
    // the minimum does not depend on the order: k_dt_final reads up to
    // RED_FUSE chunks of per-tile minima itself
    if (nmin > (unsigned)(RED_FUSE * RED_CHUNK)) {
      const int nb1 = (int)((nmin + RED_CHUNK - 1) / RED_CHUNK);
      k_reduce<true><<<nb1, 256, 0, c->stream>>>(L.pmin.p, (int)nmin, L.red2.p, halt);
      k_dt_final<<<1, 256, 0, c->stream>>>(L.red2.p, nb1, c->scal.p, L.clk.p, L.cnt.p, p->time_accurate);
    } else {
      k_dt_final<<<1, 256, 0, c->stream>>>(L.pmin.p, (int)nmin, c->scal.p, L.clk.p, L.cnt.p, p->time_accurate);
    }
    prof_end(c, pid);
    HIPCHK(hipGetLastError());
    return 0;
  };
  // (sdt: after the first stage, which reads neither dtglobal nor the clock)
  if (!sdt && !fdt) {
    const int e = dt_final();
    if (e) return e;
  }
  L.dt_next = 0;
  const int o = L.cur, a1 = (L.cur + 1) % 3, a2 = (L.cur + 2) % 3;
  R* in = V.Tb[o];
  int npart = (int)ntile;  // partial sums written by the kernel that applies the last update
  L.plist = 0;
  for (int irk = 1; irk <= nrk; irk++) {
    R* out = (irk % 2 == 1) ? V.Tb[a1] : V.Tb[a2];
    // Fill_Ghost_Cell: only read by tangential gradients (non-orthogonal
    // blocks; 300/302 walls) and connections; a level with neither skips it
    // (fused stages write the ghosts of their outputs; the step's input gets
    // its ghosts here)
    const bool gh = L.tang || L.ghost_read || c->adj_store_on;
    int e = gh && (!fused || irk == 1) ? launch_ghost(c, L, in) : 0;
    if (e) return e;
    if constexpr (std::is_same<R, double>::value) {
      if (c->adj_store_on)
        HIPCHK(hipMemcpyAsync(c->adj.stage[irk - 1].p, in, L.npad * sizeof(double), hipMemcpyDeviceToDevice,
                              c->stream));
    }
    StageArgs<R> sa{};
    sa.blk = L.blk.p; sa.tiles = L.tiles.p; sa.T = in; sa.TO = V.Tb[o]; sa.Tout = out; sa.Rres = V.res;
    sa.dt = V.dt; sa.dtg = c->scal.p; sa.use_dtg = p->time_accurate; sa.qvol = V.qvol; sa.vol = V.vol;
    sa.mat = L.mat.p;
    for (int w = 0; w < 9; w++) sa.W[w] = V.W[w];
    sa.bc = L.bc.p; sa.fmap = L.fmap.p; sa.bcv = V.bcv; sa.bcf = V.bcf; sa.tab = dtab<R>(c);
    sa.bcz = (unsigned)(L.bcf.n - 1);
    sa.ser = serdev(c); sa.clk = L.clk.p; sa.halt = halt;
    sa.partials = L.partials.p; sa.status = c->status.p;  // k_bcflux flags table errors through status
    sa.rect = (L.rect_w && !c->rect_off) ? L.rfac.p : nullptr;
    if (sdt && irk == 1) {
      sa.dtw = V.dt; sa.vnn = p->vnn; sa.pmin = L.pmin.p;
    }
    if (L.any_active && L.nbound && !fused) {
      pid = prof_begin(c, FUSS_K_BC);
      sa.conn_split = L.nconn > 0;
      if (L.kp == GC) k_bcflux<3, R><<<(L.nbound + 127) / 128, 128, 0, c->stream>>>(sa, L.bch.p, L.nbound);
      else k_bcflux<2, R><<<(L.nbound + 127) / 128, 128, 0, c->stream>>>(sa, L.bch.p, L.nbound);
      if (L.nconn) {
        if (L.kp == GC) k_connflux<3, R><<<(L.nconn + 127) / 128, 128, 0, c->stream>>>(sa, L.cd.p, L.ci.p, L.nconn);
        else k_connflux<2, R><<<(L.nconn + 127) / 128, 128, 0, c->stream>>>(sa, L.cd.p, L.ci.p, L.nconn);
      }
      sa.conn_split = 0;
      prof_end(c, pid);
      HIPCHK(hipGetLastError());
    }
    sa.rkc = nrk == 4 ? RK4A[irk - 1] : RKC[nrk - 1][irk - 1];
    sa.jam = nrk == 4;
    sa.last = irk == nrk && p->compute_residual;
    sa.irk = irk;
    sa.ctiles = L.ctiles.p;
    pid = prof_begin(c, (sdt && irk == 1) ? FUSS_K_STAGE_DT : FUSS_K_STAGE);
    if (hk) {
      // Fluxes + BC_Fluxes to R, External_Function on the host, RK_Newstate
      // (Mod_Explicit.f90:53-59)
      sa.Rres = V.res;
      const unsigned gt = (unsigned)L.htiles.size();
      launch_stage_raw(c, L, sa, c->ndir, L.tang != 0);
      HIPCHK(hipGetLastError());
      int eh = FUSS_EUNSUPPORTED;  // (FP64 steps only)
      if constexpr (std::is_same<R, double>::value) eh = run_hook(c, L, level, irk, in, hk);
      if (eh) return eh;
      launch_newstate(c, L, sa, prim, irs);
      if (!irs) { npart = (int)gt; L.plist = 0; }
    } else if (fused) {
      DtArgs<R> da{};
      da.blk = L.blk.p; da.tiles = L.tiles.p; da.T = in; da.mat = L.mat.p; da.invL2 = V.invL2; da.dt = V.dt;
      da.tab = dtab<R>(c); da.vnn = p->vnn; da.cnt = L.cnt.p; da.rampa = p->rampa_iter; da.pmin = L.pmin.p;
      da.status = c->status.p;
      launch_stage_fused(c, L, sa, da, c->ndir, L.tang != 0, prim, fdt && irk == 1, gh);
      npart = (int)ntile;
      L.plist = 0;
    } else if (L.kp == GC) {
      launch_stage3d(c, L, sa, L.tang != 0, prim, irs);
      if (!irs) { npart = (int)nstcol; L.plist = stlist; }
    } else {
      launch_stage2d(c, L, sa, L.tang != 0, prim, irs);
    }
    prof_end(c, pid);
    HIPCHK(hipGetLastError());
    if (sdt && irk == 1) {
      const int e2 = dt_final();
      if (e2) return e2;
    }
    if (irs) {
      // Residual_Smoothing (Lib_IRS.f90:14-40): ndir passes R -> R2 -> R ...
      IrsArgs<R> ia{};
      memset(&ia, 0, sizeof(ia));
      ia.blk = L.blk.p; ia.tiles = L.tiles.p; ia.bc = L.bc.p; ia.fmap = L.fmap.p;
      ia.beta = p->irs_beta;
      {
#pragma clang fp contract(off)
        ia.den = 1.0 + 2.0 * p->irs_beta;
        ia.inv = exact_inverse(ia.den);
      }
      ia.T = in; ia.TO = V.Tb[o]; ia.Tout = out; ia.mat = L.mat.p; ia.tab = dtab<R>(c);
      ia.rkc = sa.rkc; ia.jam = sa.jam; ia.last = sa.last; ia.irk = irk; ia.partials = L.partials.p; ia.status = c->status.p;
      ia.halt = halt;
      if (tri) {
        int np_ = 0;
        const int e = launch_irs_tridiag(c, L, ia, V.res, V.res2, prim, &np_);
        if (e) return e;
        if (sa.last) { npart = np_; L.plist = 0; }
      } else if (xyz) {
        // x, y and z in one k-marching pass with the update (and on the last
        // stage the next step's dt)
        const bool c8 = L.xyz_ty == 8;
        ia.Rin = V.res;
        ia.tiles = c8 ? L.c8tiles.p : L.ctiles.p;
        ia.dtnext = fuse_dt && irk == nrk;
        ia.vnn = p->vnn; ia.invL2 = V.invL2; ia.dt = V.dt; ia.pmin = L.pmin.p;
        ia.rect = (L.rect_dt && !c->rect_off) ? L.rfac.p : nullptr;
        pid = prof_begin(c, FUSS_K_IRS_UPDATE);
        launch_irs_xyz(c, L, ia, prim);
        prof_end(c, pid);
        HIPCHK(hipGetLastError());
        if (ia.dtnext) L.dt_ty = L.xyz_ty;
        if (sa.last) {
          npart = (int)(c8 ? L.hc8tiles.size() : L.hctiles.size());
          L.plist = c8 ? 3 : 1;
        }
      } else if (L.no_ycon && c->ndir >= 2) {
        // x and y in one LDS pass; 2-D levels update there, 3-D levels then
        // march the z pass with the update
        ia.Rin = V.res;
        ia.Rout = V.res2;
        ia.tiles = L.xytiles.p;
        const unsigned gxy = (unsigned)L.hxytiles.size();
        pid = prof_begin(c, c->ndir == 2 ? FUSS_K_IRS_UPDATE : FUSS_K_IRS);
        launch_irs_xy(c, L, c->ndir, ia, prim);
        if (c->ndir == 2 && sa.last) { npart = (int)gxy; L.plist = 2; }
        prof_end(c, pid);
        HIPCHK(hipGetLastError());
        if (c->ndir == 3) {
          ia.Rin = V.res2;
          ia.tiles = L.ctiles.p;
          pid = prof_begin(c, FUSS_K_IRS_UPDATE);
          launch_irs_z(c, L, ia, prim);
          prof_end(c, pid);
          HIPCHK(hipGetLastError());
          if (sa.last) { npart = (int)L.hctiles.size(); L.plist = 1; }
        }
      } else {
        R* rin = V.res;
        R* rout = V.res2;
        const int nline = c->ndir == 3 ? 2 : c->ndir;  // 3-D: z pass marched below
        for (int d = 0; d < nline; d++) {
          ia.Rin = rin; ia.Rout = rout;
          const bool upd = d == c->ndir - 1;
          pid = prof_begin(c, upd ? FUSS_K_IRS_UPDATE : FUSS_K_IRS);
          launch_irs_line(c, L, c->ndir, d, ia, upd, prim);
          prof_end(c, pid);
          HIPCHK(hipGetLastError());
          std::swap(rin, rout);
        }
        if (c->ndir == 3) {
          ia.Rin = rin;
          ia.tiles = L.ctiles.p;
          pid = prof_begin(c, FUSS_K_IRS_UPDATE);
          launch_irs_z(c, L, ia, prim);
          prof_end(c, pid);
          HIPCHK(hipGetLastError());
          if (sa.last) { npart = (int)L.hctiles.size(); L.plist = 1; }
        }
      }
    }
    in = out;
  }
  if (fused && (p->compute_residual || fdt)) {
    // the residual sum (and, when the first stage formed dt, its minimum and
    // the clock) in one work-group
    pid = prof_begin(c, FUSS_K_REDUCE);
    launch_fused_end(c, L, (int)ntile, fdt, p->time_accurate);
    prof_end(c, pid);
    HIPCHK(hipGetLastError());
  } else if (p->compute_residual) {
    pid = prof_begin(c, FUSS_K_REDUCE);
    const int nb1 = (npart + RED_CHUNK - 1) / RED_CHUNK;
    if (nb1 <= RED_FUSE) {
      k_reduce_sum1<<<1, 256, 0, c->stream>>>(L.partials.p, npart, c->scal.p + 2, halt);
    } else {
      k_reduce<false><<<nb1, 256, 0, c->stream>>>(L.partials.p, npart, L.red2.p, halt);
      k_reduce<false><<<1, 256, 0, c->stream>>>(L.red2.p, nb1, c->scal.p + 2, halt);
    }
    prof_end(c, pid);
    HIPCHK(hipGetLastError());
  }
  // T^{n+1} becomes the new T^n
  L.cur = (nrk % 2 == 1) ? a1 : a2;
  L.dt_next = fuse_dt;
  L.dt_vnn = p->vnn;
  return 0;
}

// Explicit_Step in the context's precision: FP32 contexts step their FP32
// mirrors (the state is converted once, before the first FP32 step, outside
// any capture); hooks, adjoint recording and chimera ghosts are FP64 only.
int enqueue_step(fuss_ctx* c, int level, const fuss_step_params* p, const StageHook* hk = nullptr) {
  Level& L = c->lev[level - 1];
  if (c->precision == FUSS_FP32) {
    if (hk || c->adj_store_on || L.nchim) return FUSS_EUNSUPPORTED;
    const int e = state_to_f32(c, L);
    if (e) return e;
    return enqueue_step_t<float>(c, level, p, hk);
  }
  const int e = state_to_f64(c, L);
  if (e) return e;
  return enqueue_step_t<double>(c, level, p, hk);
}

// FP64 step whatever the context's precision (the defect correction's
// monitor, the adjoint's recorded forward steps)
int enqueue_step_f64(fuss_ctx* c, int level, const fuss_step_params* p) {
  const int e = state_to_f64(c, c->lev[level - 1]);
  if (e) return e;
  return enqueue_step_t<double>(c, level, p, nullptr);
}

// ------------------------------------------------------------------------
// steady defect correction (fuss_refine)
// ------------------------------------------------------------------------

// Stage coefficients a_1..a_m of the multistage smoother whose polynomial is
// the Chebyshev one on [lo, hi]: P(z) = prod_i (1 + z / l_i)
// = 1 + a_m z (1 + a_{m-1} z (... (1 + a_1 z))), z = -lambda.
std::vector<double> cheb_stages(int m, double lo, double hi) {
  std::vector<double> cf(1, 1.0);
  for (int i = 1; i <= m; i++) {
    const double l = 0.5 * (hi + lo) + 0.5 * (hi - lo) * cos(M_PI * (2 * i - 1) / (2.0 * m));
    std::vector<double> nc(cf.size() + 1, 0.0);
    for (size_t q = 0; q < cf.size(); q++) {
      nc[q] += cf[q];
      nc[q + 1] += cf[q] / l;
    }
    cf.swap(nc);
  }
  std::vector<double> a(m);
  for (int q = 1; q <= m; q++) a[m - q] = cf[q] / cf[q - 1];
  return a;
}

void refine_free(Refine& R) {
  for (auto& l : R.lv) {
    l.blk.free();
    l.mem.free();
  }
  R.lv.clear();
  R.bcd.free();
  R.red.free();
  R.bres.free();
  if (R.hbres) (void)hipHostFree(R.hbres);
  R.hbres = nullptr;
  R.ready = 0;
}

inline unsigned cgrid(int n) { return (unsigned)((n + 255) / 256); }

int refine_build(fuss_ctx* c, const fuss_refine_params* rp) {
  Refine& R = c->ref;
  Level& L = c->lev[0];
  for (int e = 0; e < L.nbound; e++) {
    const int t = L.hib[15 * e + 5];
    if (t == 101 || t == 102) return FUSS_EUNSUPPORTED;
  }
  if (rp->precision != FUSS_FP32 && rp->precision != FUSS_FP64) return FUSS_EARG;
  if (rp->n_stage < 1 || rp->n_stage > 8 || rp->nu1 < 0 || rp->nu2 < 0 || rp->vcycles < 1 || rp->max_outer < 0 ||
      rp->max_levels < 0 || rp->ncoarse < 0 || rp->coarse_cg < 0 || !(rp->lam_lo > 0.0) ||
      !(rp->lam_hi > rp->lam_lo) || !(rp->irs_beta >= 0.0))
    return FUSS_EARG;
  refine_free(R);
  const size_t es = rp->precision == FUSS_FP32 ? 4 : 8;
  const int nb = L.nb;
  std::vector<CBlk> prev;
  for (int l = 0;; l++) {
    CLevel cl;
    cl.hblk.resize(nb);
    long off = 0;
    bool coarsened = false;
    for (int b = 0; b < nb; b++) {
      CBlk& B = cl.hblk[b];
      B = CBlk{};
      const int fn[3] = {l ? prev[b].n[0] : L.hblk[b].ni, l ? prev[b].n[1] : L.hblk[b].nj,
                         l ? prev[b].n[2] : L.hblk[b].nk};
      for (int d = 0; d < 3; d++) {
        B.r[d] = (l > 0 && fn[d] > 2) ? 2 : 1;
        if (B.r[d] == 2) coarsened = true;
        B.n[d] = (fn[d] + B.r[d] - 1) / B.r[d];
        B.nf[d] = fn[d];
      }
      B.off = (int)off;
      B.foff = l ? prev[b].off : 0;
      const long nc = (long)B.n[0] * B.n[1] * B.n[2];
      cl.maxcells = (int)std::max<long>(cl.maxcells, nc);
      off += nc;
    }
    if (l > 0 && !coarsened) break;
    if (off >= (1L << 31) / 8) return FUSS_EUNSUPPORTED;
    cl.ncell = (int)off;
    prev = cl.hblk;
    R.lv.push_back(std::move(cl));
    if (rp->max_levels > 0 && (int)R.lv.size() >= rp->max_levels) break;
  }
  for (auto& cl : R.lv) {
    const size_t one = ((size_t)cl.ncell * es + 255) / 256 * 256;
    if (cl.mem.alloc(10 * one)) return FUSS_EHIP;
    unsigned char* base = cl.mem.p;
    void** slots[10] = {&cl.d[0], &cl.d[1], &cl.d[2], &cl.b, &cl.cx, &cl.cy, &cl.cz, &cl.D, &cl.s, &cl.s2};
    for (int q = 0; q < 10; q++) *slots[q] = base + q * one;
    if (cl.blk.alloc(cl.hblk.size())) return FUSS_EHIP;
    HIPCHK(hipMemcpy(cl.blk.p, cl.hblk.data(), cl.hblk.size() * sizeof(CBlk), hipMemcpyHostToDevice));
    cl.cur = 0;
  }
  {
    // conjugate gradients on the coarsest level when it is not solved directly
    CLevel& cl = R.lv.back();
    if (rp->coarse_cg > 0 && cl.maxcells > CMG_DIRECT) {
      const size_t one = ((size_t)cl.ncell * es + 255) / 256 * 256;
      if (cl.cgmem.alloc(6 * one) || cl.cgpart.alloc(3 * (size_t)cgrid(cl.ncell) + 2)) return FUSS_EHIP;
      for (int q = 0; q < 6; q++) cl.cg[q] = cl.cgmem.p + q * one;
    }
  }
  if (R.bcd.alloc(std::max(L.nbound, 1)) || R.red.alloc(2) || R.bres.alloc(nb) ||
      hipHostMalloc((void**)&R.hbres, nb * sizeof(double), 0) != hipSuccess)
    return FUSS_EHIP;
  R.p = *rp;
  R.ready = 1;
  return 0;
}

template <typename S>
COp<S> cop(const CLevel& l) {
  COp<S> o;
  o.blk = l.blk.p;
  o.nb = (int)l.hblk.size();
  o.ncell = l.ncell;
  o.cx = (const S*)l.cx;
  o.cy = (const S*)l.cy;
  o.cz = (const S*)l.cz;
  o.D = (const S*)l.D;
  o.b = (const S*)l.b;
  return o;
}


// nu iterations of the multistage smoother on one correction level.
template <typename S>
int cmg_smooth(fuss_ctx* c, CLevel& l, int nu, const std::vector<double>& al, double beta) {
  const COp<S> o = cop<S>(l);
  const unsigned gs = cgrid(l.ncell);
  const S bt = (S)beta;
  S den;
  {
#pragma clang fp contract(off)
    den = (S)(1.0 + 2.0 * beta);
  }
  const S inv = (S)exact_inverse((double)den);
  const int m = (int)al.size();
  for (int it = 0; it < nu; it++) {
    const int base = l.cur, x = (base + 1) % 3, y = (base + 2) % 3;
    const S* d0 = (const S*)l.d[base];
    const S* din = d0;
    for (int st = 0; st < m; st++) {
      S* out = (S*)l.d[st % 2 == 0 ? x : y];
      const S a = (S)al[st];
      if (beta <= 0.0) {
        k_cmg_stage<S, false><<<gs, 256, 0, c->stream>>>(o, din, d0, out, a);
      } else {
        S* s1 = (S*)l.s;
        S* s2 = (S*)l.s2;
        k_cmg_stage<S, true><<<gs, 256, 0, c->stream>>>(o, din, nullptr, s1, (S)0);
        if (c->ndir == 3) {
          k_cmg_irs<0, false, S><<<gs, 256, 0, c->stream>>>(o, s1, s2, bt, den, inv, nullptr, (S)0);
          k_cmg_irs<1, false, S><<<gs, 256, 0, c->stream>>>(o, s2, s1, bt, den, inv, nullptr, (S)0);
          k_cmg_irs<2, true, S><<<gs, 256, 0, c->stream>>>(o, s1, out, bt, den, inv, d0, a);
        } else {
          k_cmg_irs<0, false, S><<<gs, 256, 0, c->stream>>>(o, s1, s2, bt, den, inv, nullptr, (S)0);
          k_cmg_irs<1, true, S><<<gs, 256, 0, c->stream>>>(o, s2, out, bt, den, inv, d0, a);
        }
      }
      HIPCHK(hipGetLastError());
      din = out;
    }
    l.cur = (m % 2 == 1) ? x : y;
  }
  return 0;
}

// iters conjugate-gradient iterations on a coarsest level from its current
// correction (Jacobi preconditioner; k_cg_*)
template <typename S>
int cmg_cg(fuss_ctx* c, CLevel& l, int iters) {
  const COp<S> o = cop<S>(l);
  const unsigned gs = cgrid(l.ncell);
  S* x = (S*)l.d[l.cur];
  S *r = (S*)l.cg[0], *z = (S*)l.cg[1], *p0 = (S*)l.cg[2], *p1 = (S*)l.cg[3], *q = (S*)l.cg[4];
  S* invd = (S*)l.cg[5];
  // partials: r . z (two, alternating by iteration), p . q; the last two r . z
  double* prz[2] = {l.cgpart.p, l.cgpart.p + gs};
  double* ppq = l.cgpart.p + 2 * gs;
  double* rzs = l.cgpart.p + 3 * gs;
  k_cg_start<S><<<gs, 256, 0, c->stream>>>(o, x, r, z, p0, invd, prz[0]);
  HIPCHK(hipGetLastError());
  for (int it = 0; it < iters; it++) {
    // p' = z + beta p alternates between the two direction buffers
    const S* pold = it % 2 == 0 ? p0 : p1;
    S* pnew = it % 2 == 0 ? p1 : p0;
    k_cg_ap<S><<<gs, 256, 0, c->stream>>>(o, it, z, pold, pnew, q, prz[it % 2], ppq, rzs);
    k_cg_update<S><<<gs, 256, 0, c->stream>>>(l.ncell, x, pnew, r, q, invd, z, prz[it % 2], ppq, prz[(it + 1) % 2]);
    HIPCHK(hipGetLastError());
  }
  return 0;
}

template <typename S>
int cmg_vcycle(fuss_ctx* c, int l, const std::vector<double>& al) {
  Refine& R = c->ref;
  CLevel& L = R.lv[l];
  const int nl = (int)R.lv.size();
  if (l == nl - 1) {
    if (L.maxcells <= CMG_DIRECT) {
      const int nb = (int)L.hblk.size();
      k_cmg_direct<S><<<(nb + 63) / 64, 64, 0, c->stream>>>(cop<S>(L), (S*)L.d[L.cur]);
      HIPCHK(hipGetLastError());
      return 0;
    }
    if (R.p.coarse_cg > 0) return cmg_cg<S>(c, L, R.p.coarse_cg);
    return cmg_smooth<S>(c, L, R.p.ncoarse, al, R.p.irs_beta);
  }
  int e = cmg_smooth<S>(c, L, R.p.nu1, al, R.p.irs_beta);
  if (e) return e;
  CLevel& C = R.lv[l + 1];
  k_cmg_restrict<S><<<cgrid(C.ncell), 256, 0, c->stream>>>(cop<S>(L), (const S*)L.d[L.cur], cop<S>(C), (S*)C.b,
                                                          (S*)C.d[C.cur]);
  HIPCHK(hipGetLastError());
  e = cmg_vcycle<S>(c, l + 1, al);
  if (e) return e;
  k_cmg_prolong<S><<<cgrid(L.ncell), 256, 0, c->stream>>>(cop<S>(L), (S*)L.d[L.cur], cop<S>(C),
                                                         (const S*)C.d[C.cur]);
  HIPCHK(hipGetLastError());
  return cmg_smooth<S>(c, L, R.p.nu2, al, R.p.irs_beta);
}

template <int NDIR, bool TANG, typename S>
void launch_defect(fuss_ctx* c, Level& L, const StageArgs<double>& sa, CLevel& C0) {
  k_defect<NDIR, TANG, S><<<(unsigned)L.htiles.size(), TILE_T, 0, c->stream>>>(
      sa, c->ref.bcd.p, C0.blk.p, (S*)C0.b, (S*)C0.cx, (S*)C0.cy, (S*)C0.cz, (S*)C0.D, L.partials.p);
}

template <typename S>
int refine_run(fuss_ctx* c, const fuss_step_params* mon, fuss_refine_result* out, double* hist) {
  Refine& R = c->ref;
  Level& L = c->lev[0];
  const fuss_refine_params& rp = R.p;
  const std::vector<double> al = cheb_stages(rp.n_stage, rp.lam_lo, rp.lam_hi);
  fuss_step_params pm = *mon;
  pm.compute_residual = 1;
  CLevel& C0 = R.lv[0];
  const int nl = (int)R.lv.size();
  const unsigned ntile = (unsigned)L.htiles.size();
  *out = fuss_refine_result{};
  out->levels = nl;
  R.conv_outer.assign(L.nb, -1);
  R.res.assign(L.nb, 0.0);
  bool have_def = false;
  for (int m = 0;; m++) {
    // monitor: one FP64 Explicit_Step; its residual is the reference's
    drop_dt_next(c);  // T was corrected since the last monitor step
    int e = enqueue_step_f64(c, 1, &pm);
    if (e) return e;
    const int nb = L.nb;
    k_block_sums<<<nb, 256, 0, c->stream>>>(L.partials.p, L.rng[L.plist].p, R.bres.p);
    HIPCHK(hipGetLastError());
    HIPCHK(hipMemcpyAsync(c->hscal, c->scal.p, 3 * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipMemcpyAsync(R.hbres, R.bres.p, nb * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    if (have_def) HIPCHK(hipMemcpyAsync(c->hscal + 4, R.red.p, sizeof(double), hipMemcpyDeviceToHost, c->stream));
    int bijk[4] = {0, 0, 0, 0};
    const int st = sync_status(c, bijk);
    if (st) return st;
    const double res = sqrt(c->hscal[2]);
    if (have_def) out->defect = sqrt(c->hscal[4]);
    if (hist) hist[m] = res;
    out->res = res;
    out->outer = m;
    int all = 1;
    for (int b = 0; b < nb; b++) {
      R.res[b] = sqrt(R.hbres[b]);
      if (R.res[b] <= rp.res_threshold) {
        if (R.conv_outer[b] < 0) R.conv_outer[b] = m;
      } else {
        all = 0;
      }
    }
    // one domain: the reference's residual over all blocks; independent
    // cases (per_block): every block's own residual
    if (rp.per_block ? all : res <= rp.res_threshold) {
      out->converged = 1;
      break;
    }
    if (m >= rp.max_outer) break;
    // FP64 defect, conductances and BC diagonal at the current T
    double* T = L.Tb[L.cur].p;
    e = launch_ghost(c, L, T);
    if (e) return e;
    StageArgs<double> sa{};
    memset(&sa, 0, sizeof(sa));
    sa.blk = L.blk.p; sa.tiles = L.tiles.p; sa.ctiles = L.ctiles.p; sa.T = T; sa.TO = T;
    sa.qvol = L.qvol.p; sa.vol = L.vol.p; sa.mat = L.mat.p;
    for (int w = 0; w < 9; w++) sa.W[w] = L.W[w].p;
    sa.bc = L.bc.p; sa.fmap = L.fmap.p; sa.bcv = L.bcv.p; sa.bcf = L.bcf.p; sa.tab = dtab<double>(c);
    sa.ser = serdev(c); sa.clk = L.clk.p; sa.halt = L.cnt.p + 1;
    sa.partials = L.partials.p; sa.status = c->status.p;
    const bool d3 = L.kp == GC;
    if (L.nbound) {
      if (L.any_active) {
        d3 ? k_bcflux<3, double><<<(L.nbound + 127) / 128, 128, 0, c->stream>>>(sa, L.bch.p, L.nbound)
           : k_bcflux<2, double><<<(L.nbound + 127) / 128, 128, 0, c->stream>>>(sa, L.bch.p, L.nbound);
      }
      d3 ? k_bcderiv<3><<<(L.nbound + 127) / 128, 128, 0, c->stream>>>(sa, R.bcd.p, L.nbound)
         : k_bcderiv<2><<<(L.nbound + 127) / 128, 128, 0, c->stream>>>(sa, R.bcd.p, L.nbound);
      HIPCHK(hipGetLastError());
    }
    if (d3) L.tang ? launch_defect<3, true, S>(c, L, sa, C0) : launch_defect<3, false, S>(c, L, sa, C0);
    else L.tang ? launch_defect<2, true, S>(c, L, sa, C0) : launch_defect<2, false, S>(c, L, sa, C0);
    HIPCHK(hipGetLastError());
    {
      const int nb1 = (int)((ntile + RED_CHUNK - 1) / RED_CHUNK);
      k_reduce<false><<<nb1, 256, 0, c->stream>>>(L.partials.p, (int)ntile, L.red2.p);
      k_reduce<false><<<1, 256, 0, c->stream>>>(L.red2.p, nb1, R.red.p);
      HIPCHK(hipGetLastError());
    }
    have_def = true;
    for (int l = 1; l < nl; l++) {
      CLevel& Cl = R.lv[l];
      k_cmg_coarsen<S><<<cgrid(Cl.ncell), 256, 0, c->stream>>>(cop<S>(R.lv[l - 1]), Cl.blk.p, (int)Cl.hblk.size(),
                                                              Cl.ncell, (S*)Cl.cx, (S*)Cl.cy, (S*)Cl.cz, (S*)Cl.D);
      HIPCHK(hipGetLastError());
    }
    k_cmg_fill<S><<<cgrid(C0.ncell), 256, 0, c->stream>>>((S*)C0.d[C0.cur], C0.ncell, (S)0);
    HIPCHK(hipGetLastError());
    for (int v = 0; v < rp.vcycles; v++) {
      e = cmg_vcycle<S>(c, 0, al);
      if (e) return e;
      out->vcycles++;
    }
    k_cmg_apply<S><<<ntile, TILE_T, 0, c->stream>>>(L.blk.p, L.tiles.p, C0.blk.p, (const S*)C0.d[C0.cur], T);
    HIPCHK(hipGetLastError());
  }
  return 0;
}


// ------------------------------------------------------------------------
// discrete adjoint of the time-accurate step (fuss_adj_*)
// ------------------------------------------------------------------------
void adj_free(Adjoint& A) {
  A.probe.free(); A.coff.free(); A.ckpt.free(); A.seg.free();
  for (auto& b : A.stage) b.free();
  A.lam.free(); A.part.free(); A.red.free(); A.acc.free(); A.amin.free(); A.dtimes.free(); A.diters.free();
  A.ready = 0;
  A.forward_done = 0;
}

int adj_setup(fuss_ctx* c, int nsteps, int K, int nprobe, const int* bijk, int prec) {
  Adjoint& A = c->adj;
  Level& L = c->lev[0];
  if (nsteps < 1 || K < 1 || nprobe < 0 || (nprobe && !bijk)) return FUSS_EARG;
  if (prec != FUSS_FP64 && prec != FUSS_FP32) return FUSS_EARG;
  if (L.tang) return FUSS_EUNSUPPORTED;
  for (int e = 0; e < L.nbound; e++) {
    const int t = L.hib[15 * e + 5];
    if (!(t == 0 || t == 301 || t == 302 || t == 303)) return FUSS_EUNSUPPORTED;
    if (t == 302 && L.hser.size() > (size_t)e && L.hser[e] >= 0) return FUSS_EUNSUPPORTED;
  }
  adj_free(A);
  A.nsteps = nsteps; A.K = K; A.nprobe = nprobe; A.prec = prec;
  A.hprobe.resize(nprobe);
  for (int q = 0; q < nprobe; q++) {
    const int b = bijk[4 * q] - 1, i = bijk[4 * q + 1], j = bijk[4 * q + 2], k = bijk[4 * q + 3];
    if (b < 0 || b >= L.nb) return FUSS_EARG;
    const BlockInfo& B = L.hblk[b];
    if (i < 1 || i > B.ni || j < 1 || j > B.nj || k < 1 || k > B.nk) return FUSS_EARG;
    A.hprobe[q] = cidx(B, i, j, k);
  }
  std::vector<long> coff(L.nb);
  long o = 0;
  for (int b = 0; b < L.nb; b++) { coff[b] = o; o += (long)L.hblk[b].ni * L.hblk[b].nj * L.hblk[b].nk; }
  const size_t np = (size_t)L.npad;
  const size_t es = prec == FUSS_FP32 ? 4 : 8;
  const size_t one = (np * es + 255) / 256 * 256;
  const size_t ntile = L.htiles.size();
  int e = A.probe.alloc(std::max(nprobe, 1)) | A.coff.alloc(L.nb) | A.ckpt.alloc(((nsteps - 1) / K + 2) * np) |
          A.seg.alloc((size_t)K * np) | A.lam.alloc(3 * one) | A.part.alloc(2 * ntile) |
          A.red.alloc(2 * (ntile / RED_CHUNK + 2)) | A.acc.alloc(2) | A.amin.alloc(ntile) |
          A.dtimes.alloc(nsteps) | A.diters.alloc(nsteps);
  for (auto& b : A.stage) e |= b.alloc(np);
  if (e) { adj_free(A); return FUSS_EHIP; }
  if (nprobe) HIPCHK(hipMemcpy(A.probe.p, A.hprobe.data(), nprobe * sizeof(long), hipMemcpyHostToDevice));
  HIPCHK(hipMemcpy(A.coff.p, coff.data(), L.nb * sizeof(long), hipMemcpyHostToDevice));
  A.ready = 1;
  return 0;
}

int adj_forward(fuss_ctx* c, const fuss_step_params* p, double* hist) {
  Adjoint& A = c->adj;
  Level& L = c->lev[0];
  if (p->integration != FUSS_PRIM || p->irs_beta > 0.0 || !p->time_accurate || p->n_rk < 1 || p->n_rk > 4)
    return FUSS_EUNSUPPORTED;
  A.p = *p;
  A.p.compute_residual = 0;
  A.times.assign(A.nsteps, 0.0);
  A.iters.assign(A.nsteps, 0);
  const size_t np = (size_t)L.npad;
  for (int n = 0; n < A.nsteps; n++) {
    if (n % A.K == 0)
      HIPCHK(hipMemcpyAsync(A.ckpt.p + (size_t)(n / A.K) * np, L.Tb[L.cur].p, np * sizeof(double),
                            hipMemcpyDeviceToDevice, c->stream));
    k_clock_rec<<<1, 64, 0, c->stream>>>(L.clk.p, L.cnt.p, A.dtimes.p + n, A.diters.p + n);
    const int e = enqueue_step_f64(c, 1, &A.p);
    if (e) return e;
    if (A.nprobe && hist) {
      k_probe_rec<<<(A.nprobe + 63) / 64, 64, 0, c->stream>>>(L.Tb[L.cur].p, A.probe.p, hist + (size_t)n * A.nprobe,
                                                            A.nprobe);
      HIPCHK(hipGetLastError());
    }
  }
  // keep T^N for restoring the level after a reverse sweep
  HIPCHK(hipMemcpyAsync(A.ckpt.p + (size_t)((A.nsteps - 1) / A.K + 1) * np, L.Tb[L.cur].p, np * sizeof(double),
                        hipMemcpyDeviceToDevice, c->stream));
  const int st = sync_status(c);
  if (st) return st;
  HIPCHK(hipMemcpy(A.times.data(), A.dtimes.p, A.nsteps * sizeof(double), hipMemcpyDeviceToHost));
  HIPCHK(hipMemcpy(A.iters.data(), A.diters.p, A.nsteps * sizeof(int), hipMemcpyDeviceToHost));
  if (read_clock(c, L)) return FUSS_EHIP;
  A.forward_done = 1;
  return 0;
}

template <int NDIR, typename S>
int adj_backward_t(fuss_ctx* c, const double* g, double* gqw, double* gk, double* gT0) {
  Adjoint& A = c->adj;
  Level& L = c->lev[0];
  const size_t np = (size_t)L.npad;
  const size_t one = (np * sizeof(S) + 255) / 256 * 256;
  S* lam = (S*)A.lam.p;
  S* lin = (S*)(A.lam.p + one);
  S* lTO = (S*)(A.lam.p + 2 * one);
  const unsigned fb = (unsigned)((np + 255) / 256);
  k_adj_fill<S><<<fb, 256, 0, c->stream>>>(lam, (long)np);
  k_adj_fill<S><<<fb, 256, 0, c->stream>>>(lin, (long)np);
  k_adj_fill<S><<<fb, 256, 0, c->stream>>>(lTO, (long)np);
  k_adj_fill<double><<<1, 64, 0, c->stream>>>(A.acc.p, 2);
  if (gqw && L.nbound) k_adj_fill<double><<<(L.nbound + 255) / 256, 256, 0, c->stream>>>(gqw, L.nbound);
  HIPCHK(hipGetLastError());
  DBuf<double> gq_own;
  if (!gqw) {
    if (gq_own.alloc(std::max(L.nbound, 1))) return FUSS_EHIP;
    gqw = gq_own.p;
  }
  const unsigned ntile = (unsigned)L.htiles.size();
  const int nb1 = (int)((ntile + RED_CHUNK - 1) / RED_CHUNK);
  const int nrk = A.p.n_rk;
  const double t_end = L.time;
  const int it_end = L.iter;
  const int N = A.nsteps, K = A.K;
  if (A.nprobe && g)
    k_adj_seed<S><<<1, 64, 0, c->stream>>>(lam, A.probe.p, g + (size_t)(N - 1) * A.nprobe, A.nprobe);
  const int nseg = (N + K - 1) / K;
  int err = 0;
  for (int sg = nseg - 1; sg >= 0 && !err; sg--) {
    const int n0 = sg * K, n1 = std::min(N, n0 + K);
    HIPCHK(hipMemcpyAsync(L.Tb[L.cur].p, A.ckpt.p + (size_t)sg * np, np * sizeof(double), hipMemcpyDeviceToDevice,
                          c->stream));
    for (int n = n0; n < n1; n++) {
      HIPCHK(hipMemcpyAsync(A.seg.p + (size_t)(n - n0) * np, L.Tb[L.cur].p, np * sizeof(double),
                            hipMemcpyDeviceToDevice, c->stream));
      if (n < n1 - 1) {
        set_clock(c, L, A.times[n], A.iters[n]);
        if ((err = enqueue_step_f64(c, 1, &A.p))) break;
      }
    }
    for (int n = n1 - 1; n >= n0 && !err; n--) {
      const double* Tn = A.seg.p + (size_t)(n - n0) * np;
      HIPCHK(hipMemcpyAsync(L.Tb[L.cur].p, Tn, np * sizeof(double), hipMemcpyDeviceToDevice, c->stream));
      set_clock(c, L, A.times[n], A.iters[n]);
      c->adj_store_on = 1;
      err = enqueue_step_f64(c, 1, &A.p);
      c->adj_store_on = 0;
      if (err) break;
      for (int sidx = nrk; sidx >= 1; sidx--) {
        StageArgs<double> sa{};
        memset(&sa, 0, sizeof(sa));
        sa.blk = L.blk.p; sa.tiles = L.tiles.p; sa.T = A.stage[sidx - 1].p; sa.qvol = L.qvol.p; sa.vol = L.vol.p;
        sa.mat = L.mat.p;
        for (int w = 0; w < 9; w++) sa.W[w] = L.W[w].p;
        sa.bc = L.bc.p; sa.fmap = L.fmap.p; sa.bcv = L.bcv.p; sa.bcf = L.bcf.p; sa.tab = dtab<double>(c); sa.status = c->status.p;
        sa.ser = serdev(c); sa.clk = L.clk.p; sa.halt = L.cnt.p + 1;
        if (L.nbound && L.any_active)
          k_bcflux<NDIR, double><<<(L.nbound + 127) / 128, 128, 0, c->stream>>>(sa, L.bch.p, L.nbound);
        AdjArgs<S> aa{};
        aa.blk = L.blk.p; aa.tiles = L.tiles.p; aa.T = A.stage[sidx - 1].p; aa.qvol = L.qvol.p; aa.vol = L.vol.p;
        aa.mat = L.mat.p;
        for (int d = 0; d < 3; d++) aa.W[d] = L.W[4 * d].p;
        aa.bc = L.bc.p; aa.fmap = L.fmap.p; aa.bcv = L.bcv.p; aa.tab = dtab<double>(c); aa.dtg = c->scal.p;
        aa.coef = nrk == 4 ? RK4A[sidx - 1] : RKC[nrk - 1][sidx - 1];
        aa.jam = nrk == 4;
        aa.kscale = c->kscale;
        aa.lout = lam; aa.lin = lin; aa.lTO = lTO;
        aa.part = A.part.p; aa.gqw = gqw;
        k_adj_stage<NDIR, S><<<ntile, TILE_T, 0, c->stream>>>(aa);
        k_adj_sum2<<<nb1, 256, 0, c->stream>>>(A.part.p, (int)ntile, A.red.p);
        k_adj_acc2<<<1, 64, 0, c->stream>>>(A.red.p, nb1, A.acc.p);
        HIPCHK(hipGetLastError());
        std::swap(lam, lin);
      }
      k_adj_combine<S><<<ntile, TILE_T, 0, c->stream>>>(L.blk.p, L.tiles.p, lam, lTO);
      k_adj_argmin<<<ntile, TILE_T, 0, c->stream>>>(L.blk.p, L.tiles.p, L.dt.p, c->scal.p, A.amin.p);
      k_adj_dtpath<<<1, 64, 0, c->stream>>>(A.amin.p, (int)ntile, Tn, L.mat.p, c->scal.p, dtab<double>(c), c->kscale,
                                           sizeof(S) == 8 ? (double*)lam : nullptr,
                                           sizeof(S) == 4 ? (float*)lam : nullptr, A.acc.p);
      if (n >= 1 && A.nprobe && g)
        k_adj_seed<S><<<1, 64, 0, c->stream>>>(lam, A.probe.p, g + (size_t)(n - 1) * A.nprobe, A.nprobe);
      HIPCHK(hipGetLastError());
    }
  }
  if (!err) {
    if (gT0) k_adj_extract<S><<<ntile, TILE_T, 0, c->stream>>>(L.blk.p, L.tiles.p, A.coff.p, lam, gT0);
    if (gk) HIPCHK(hipMemcpyAsync(gk, A.acc.p + 1, sizeof(double), hipMemcpyDeviceToDevice, c->stream));
    HIPCHK(hipGetLastError());
  }
  // the level returns to T^N and its clock
  HIPCHK(hipMemcpyAsync(L.Tb[L.cur].p, A.ckpt.p + (size_t)((N - 1) / K + 1) * np, np * sizeof(double),
                        hipMemcpyDeviceToDevice, c->stream));
  set_clock(c, L, t_end, it_end);
  const int st = sync_status(c);
  gq_own.free();
  return err ? err : st;
}
}  // namespace

extern "C" {

int fuss_graphs(fuss_ctx* c, int enable) {
  if (!c) return FUSS_EARG;
  c->graphs_off = !enable;
  return 0;
}

int fuss_persistent(fuss_ctx* c, int enable) {
  if (!c) return FUSS_EARG;
  c->persist = enable != 0;
  return 0;
}

int fuss_fused(fuss_ctx* c, int enable) {
  if (!c) return FUSS_EARG;
  c->fuse_small = enable != 0;
  for (auto& L : c->lev) run_graph_free(L);  // captured launches name the kernels
  return 0;
}

int fuss_wall(fuss_ctx* c, int level, double* Tw, double* qw) {
  if (!c || !c->setup_done || level < 1 || level > c->nlev || !Tw || !qw) return FUSS_EARG;
  Level& L = c->lev[level - 1];
  if (!L.nbound) return 0;
  if (state_to_f64(c, L)) return FUSS_EHIP;
  double* T = L.Tb[L.cur].p;
  // ghosts mirrored from the current state (the 302 tangential differences
  // of face-corner cells read them)
  int e = launch_ghost(c, L, T);
  if (e) return e;
  StageArgs<double> sa{};
  sa.blk = L.blk.p; sa.T = T; sa.mat = L.mat.p; sa.bc = L.bc.p; sa.tab = dtab<double>(c);
  sa.ser = serdev(c); sa.clk = L.clk.p; sa.status = c->status.p;
  DBuf<double> out;
  if (out.alloc(2 * (size_t)L.nbound)) return FUSS_EHIP;
  if (c->ndir == 3) k_wall<3><<<(L.nbound + 127) / 128, 128, 0, c->stream>>>(sa, L.nbound, out.p, out.p + L.nbound);
  else k_wall<2><<<(L.nbound + 127) / 128, 128, 0, c->stream>>>(sa, L.nbound, out.p, out.p + L.nbound);
  HIPCHK(hipGetLastError());
  HIPCHK(hipMemcpyAsync(Tw, out.p, L.nbound * sizeof(double), hipMemcpyDeviceToHost, c->stream));
  HIPCHK(hipMemcpyAsync(qw, out.p + L.nbound, L.nbound * sizeof(double), hipMemcpyDeviceToHost, c->stream));
  const int st = sync_status(c);
  out.free();
  return st;
}

// fuss_wall with the per-GPU staging array instead of an allocation: no
// hipMalloc / hipFree or allocating call.  Every call on a GPU shares the
// array (calls that overlap on one GPU race); k_wall writes every entry, so
// nothing needs clearing first.
int fuss_wall_noalloc(fuss_ctx* c, int level, double* Tw, double* qw) {
  if (!c || !c->setup_done || level < 1 || level > c->nlev || !Tw || !qw) return FUSS_EARG;
  Level& L = c->lev[level - 1];
  if (!L.nbound) return 0;
  if (L.nbound > FUSS_WALL_NOALLOC_MAX) {
    fprintf(stderr, "fuss_wall_noalloc: %d BC entries exceed the staging capacity of %d\n", L.nbound,
            FUSS_WALL_NOALLOC_MAX);
    return FUSS_EARG;
  }
  double* const stage = (c->dev >= 0 && c->dev < MAX_DEVICES) ? g_scratch[c->dev] : nullptr;
  if (!stage) return FUSS_EHIP;
  if (state_to_f64(c, L)) return FUSS_EHIP;
  double* T = L.Tb[L.cur].p;
  int e = launch_ghost(c, L, T);
  if (e) return e;
  StageArgs<double> sa{};
  sa.blk = L.blk.p; sa.T = T; sa.mat = L.mat.p; sa.bc = L.bc.p; sa.tab = dtab<double>(c);
  sa.ser = serdev(c); sa.clk = L.clk.p; sa.status = c->status.p;
  if (c->ndir == 3) k_wall<3><<<(L.nbound + 127) / 128, 128, 0, c->stream>>>(sa, L.nbound, stage, stage + L.nbound);
  else k_wall<2><<<(L.nbound + 127) / 128, 128, 0, c->stream>>>(sa, L.nbound, stage, stage + L.nbound);
  HIPCHK(hipGetLastError());
  HIPCHK(hipMemcpyAsync(Tw, stage, L.nbound * sizeof(double), hipMemcpyDeviceToHost, c->stream));
  HIPCHK(hipMemcpyAsync(qw, stage + L.nbound, L.nbound * sizeof(double), hipMemcpyDeviceToHost, c->stream));
  return sync_status(c);
}

int fuss_rect(fuss_ctx* c, int enable) {
  if (!c) return FUSS_EARG;
  c->rect_off = !enable;
  for (auto& L : c->lev) run_graph_free(L);  // captured launches name the kernels
  return 0;
}

int fuss_rect_info(fuss_ctx* c, int level, int* weights, int* dt, double* max_rel_dev) {
  if (!c || !c->setup_done || level < 1 || level > c->nlev) return FUSS_EARG;
  const Level& L = c->lev[level - 1];
  if (weights) *weights = L.rect_w && !c->rect_off;
  if (dt) *dt = L.rect_dt && !c->rect_off;
  if (max_rel_dev) *max_rel_dev = L.rect_dev;
  return 0;
}

int fuss_set_kscale(fuss_ctx* c, double scale) {
  if (c) drop_dt_next(c);
  if (!c || !c->nT || !(scale > 0.0)) return FUSS_EARG;
  const size_t n = (size_t)c->nmat * c->nT;
  std::vector<double> k(n), k2(2 * n, 0.0);
  for (size_t q = 0; q < n; q++) k[q] = c->tkap[q] * scale;
  for (int m = 0; m < c->nmat; m++)
    for (int t = 0; t < c->nT; t++) {
      const size_t q = (size_t)m * c->nT + t;
      k2[2 * q] = k[q];
      k2[2 * q + 1] = t + 1 < c->nT ? k[q + 1] - k[q] : 0.0;
    }
  int e = upload(c->dkap.p, k.data(), n * 8) | upload(c->dkap2.p, k2.data(), 2 * n * 8);
  if (e) return FUSS_EHIP;
  c->kscale = scale;
  c->ftab_ok = 0;  // the FP32 tables follow on the next FP32 step
  for (auto& L : c->lev) run_graph_free(L);
  return 0;
}

int fuss_set_bc_values(fuss_ctx* c, int level, int n, const int* entries, int column, const double* values) {
  if (!c || !c->setup_done || level < 1 || level > c->nlev || n < 0 || (n && (!entries || !values))) return FUSS_EARG;
  if (!(column == 0 || column == 1 || column == 2 || column == 4 || column == 5)) return FUSS_EARG;
  Level& L = c->lev[level - 1];
  for (int q = 0; q < n; q++)
    if (entries[q] < 0 || entries[q] >= L.nbound) return FUSS_EARG;
  if (!n) return 0;
  for (int q = 0; q < n; q++) L.hrb[6 * entries[q] + column] = values[q];
  if ((int)c->stage_i.n < n && (c->stage_i.alloc(n) || c->stage_d.alloc(n))) return FUSS_EHIP;
  HIPCHK(hipMemcpyAsync(c->stage_i.p, entries, n * sizeof(int), hipMemcpyHostToDevice, c->stream));
  HIPCHK(hipMemcpyAsync(c->stage_d.p, values, n * sizeof(double), hipMemcpyHostToDevice, c->stream));
  k_set_bcval<<<(n + 255) / 256, 256, 0, c->stream>>>(L.bc.p, c->stage_i.p, c->stage_d.p, n, column);
  HIPCHK(hipGetLastError());
  k_bc_hot<<<(L.nbound + 255) / 256, 256, 0, c->stream>>>(L.bc.p, L.nbound, L.mat.p, L.bch.p);
  HIPCHK(hipGetLastError());
  HIPCHK(hipStreamSynchronize(c->stream));
  return 0;
}

int fuss_adj_setup(fuss_ctx* c, int nsteps, int ckpt_every, int nprobe, const int* probe_bijk, int precision) {
  if (!c || !c->setup_done) return FUSS_EARG;
  return adj_setup(c, nsteps, ckpt_every, nprobe, probe_bijk, precision);
}

int fuss_adj_forward(fuss_ctx* c, const fuss_step_params* p, double* probe_hist) {
  if (c) drop_dt_next(c);
  if (!c || !c->setup_done || !p || !c->adj.ready) return FUSS_EARG;
  if (all_to_f64(c)) return FUSS_EHIP;
  return adj_forward(c, p, probe_hist);
}

int fuss_adj_backward(fuss_ctx* c, const double* probe_adj, double* grad_qw, double* grad_kscale,
                      double* grad_T0) {
  if (c) drop_dt_next(c);
  if (!c || !c->adj.ready || !c->adj.forward_done) return FUSS_EARG;
  if (all_to_f64(c)) return FUSS_EHIP;
  const bool d3 = c->lev[0].kp == GC;
  if (c->adj.prec == FUSS_FP32)
    return d3 ? adj_backward_t<3, float>(c, probe_adj, grad_qw, grad_kscale, grad_T0)
              : adj_backward_t<2, float>(c, probe_adj, grad_qw, grad_kscale, grad_T0);
  return d3 ? adj_backward_t<3, double>(c, probe_adj, grad_qw, grad_kscale, grad_T0)
            : adj_backward_t<2, double>(c, probe_adj, grad_qw, grad_kscale, grad_T0);
}

int fuss_refine_setup(fuss_ctx* c, const fuss_refine_params* rp) {
  if (!c || !c->setup_done || !rp) return FUSS_EARG;
  if (all_to_f64(c)) return FUSS_EHIP;
  return refine_build(c, rp);
}

int fuss_refine_blocks(fuss_ctx* c, int* conv_outer, double* res) {
  if (!c || !c->ref.ready) return FUSS_EARG;
  const Refine& R = c->ref;
  for (size_t b = 0; b < R.conv_outer.size(); b++) {
    if (conv_outer) conv_outer[b] = R.conv_outer[b];
    if (res) res[b] = R.res[b];
  }
  return 0;
}

int fuss_refine(fuss_ctx* c, const fuss_step_params* monitor, fuss_refine_result* out, double* hist_res) {
  if (c) drop_dt_next(c);
  if (!c || !c->setup_done || !monitor || !out || !c->ref.ready) return FUSS_EARG;
  if (monitor->time_accurate) return FUSS_EARG;
  if (all_to_f64(c)) return FUSS_EHIP;
  const int e = c->ref.p.precision == FUSS_FP32 ? refine_run<float>(c, monitor, out, hist_res)
                                               : refine_run<double>(c, monitor, out, hist_res);
  drop_dt_next(c);
  return e;
}

}  // extern "C"

namespace {

// ------------------------------------------------------------------------
// autotuning of the 3-D column tiles (fuss_autotune)
// ------------------------------------------------------------------------
int set_kc(fuss_ctx* c, Level& L, int kc) {
  if (L.kp != GC) return 0;
  run_graph_free(L);  // its kernels address the old tile list
  L.kc = kc;
  column_tiles(L);
  if (L.ctiles.alloc(std::max<size_t>(L.hctiles.size(), 1))) return FUSS_EHIP;
  if (L.c8tiles.alloc(std::max<size_t>(L.hc8tiles.size(), 1))) return FUSS_EHIP;
  HIPCHK(hipMemcpy(L.ctiles.p, L.hctiles.data(), L.hctiles.size() * sizeof(Tile), hipMemcpyHostToDevice));
  HIPCHK(hipMemcpy(L.c8tiles.p, L.hc8tiles.data(), L.hc8tiles.size() * sizeof(Tile), hipMemcpyHostToDevice));
  HIPCHK(hipMemcpy(L.rng[1].p, L.hrng[1].data(), (L.nb + 1) * sizeof(int), hipMemcpyHostToDevice));
  HIPCHK(hipMemcpy(L.rng[3].p, L.hrng[3].data(), (L.nb + 1) * sizeof(int), hipMemcpyHostToDevice));
  const size_t np = std::max<size_t>(std::max(L.htiles.size(), L.hctiles.size()), 1);
  if (L.partials.n < np && (L.partials.alloc(np) || L.pmin.alloc(np) || L.red2.alloc(np / RED_CHUNK + 2)))
    return FUSS_EHIP;
  return 0;
}

// Times `steps` Explicit_Steps of each (column height, kernel) candidate on
// the level's current state, restores the state and clock after each, and
// keeps the fastest.
int autotune(fuss_ctx* c, int level, const fuss_step_params* p, int steps, int* kc_out, int* row_out,
             double* ms_out) {
  Level& L = c->lev[level - 1];
  if (L.kp != GC) {
    if (kc_out) *kc_out = 0;
    if (row_out) *row_out = 0;
    if (ms_out) *ms_out = 0.0;
    return 0;
  }
  if (read_clock(c, L)) return FUSS_EHIP;
  const double t0 = L.time;
  const int it0 = L.iter;
  // the FP64 state is saved and restored after each candidate (FP32
  // contexts convert it again in each candidate's untimed first step)
  if (state_to_f64(c, L)) return FUSS_EHIP;
  DBuf<double> save;
  if (save.alloc(L.npad)) return FUSS_EHIP;
  HIPCHK(hipMemcpyAsync(save.p, L.Tb[L.cur].p, L.npad * sizeof(double), hipMemcpyDeviceToDevice, c->stream));
  hipEvent_t e0, e1;
  HIPCHK(hipEventCreate(&e0));
  HIPCHK(hipEventCreate(&e1));
  fuss_step_params q = *p;
  q.compute_residual = 1;
  const int kcs[5] = {4, 8, 16, 32, 64};
  double best = 1e300;
  int bkc = L.kc, brow = 0, bst = 4, bty = 4, err = 0;
  auto trial = [&](int row, int kc, int st, int ty) -> int {
    int e = set_kc(c, L, kc);
    if (e) return e;
    L.row_march = row;
    L.stage_ty = st;
    L.xyz_ty = ty;
    drop_dt_next(c);  // the state was restored
    if ((e = enqueue_step(c, level, &q))) return e;  // warm-up
    HIPCHK(hipEventRecord(e0, c->stream));
    for (int s2 = 0; s2 < steps && !e; s2++) e = enqueue_step(c, level, &q);
    HIPCHK(hipEventRecord(e1, c->stream));
    HIPCHK(hipEventSynchronize(e1));
    float ms = 0.f;
    HIPCHK(hipEventElapsedTime(&ms, e0, e1));
    const double per = ms / std::max(steps, 1);
    if (!e && per < best) { best = per; bkc = kc; brow = row; bst = st; bty = ty; }
    // on the library's stream: the next candidate's steps follow the restore
    HIPCHK(hipMemcpyAsync(L.Tb[L.cur].p, save.p, L.npad * sizeof(double), hipMemcpyDeviceToDevice, c->stream));
    L.state_f = 0;
    set_clock(c, L, t0, it0);
    return e;
  };
  // stage kernel and column height on 4-row columns, then the column heights
  // with 8-row columns for the LDS-ring stage kernel and / or the fused x-y-z
  // IRS pass (their only users) at the chosen stage kernel
  for (int row = 0; row <= (L.tang ? 0 : 1) && !err; row++)
    for (int ki = 0; ki < 5 && !err; ki++) err = trial(row, kcs[ki], 4, 4);
  const int row4 = brow;
  const bool xyz = p->irs_beta > 0.0 && p->irs_mode == FUSS_IRS_JACOBI2 && L.no_ycon && L.no_zcon && c->ndir == 3;
  const bool ring = !L.tang && L.dims2 && row4 == 0;
  for (int st = 4; st <= (ring ? 8 : 4); st += 4)
    for (int ty = 4; ty <= (xyz ? 8 : 4); ty += 4)
      for (int ki = 0; ki < 5 && !err && (st == 8 || ty == 8); ki++) err = trial(row4, kcs[ki], st, ty);
  // the register-column stage kernel (rectilinear levels; column heights 8,
  // 16 and 32) with either IRS column
  const bool regcol = ring && L.rect_w && !c->rect_off && c->Ti1 >= 0;
  for (int ty = 4; ty <= (xyz ? 8 : 4) && regcol; ty += 4)
    for (int ki = 1; ki <= 3 && !err; ki++) err = trial(row4, kcs[ki], FUSS_STAGE_REGCOL, ty);
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  HIPCHK(hipStreamSynchronize(c->stream));  // the last restore has read `save`
  save.free();
  drop_dt_next(c);  // the state was restored after the last candidate
  if (err) return err;
  const int st = sync_status(c);
  if (st) return st;
  if ((err = set_kc(c, L, bkc))) return err;
  L.row_march = brow;
  L.stage_ty = bst;
  L.xyz_ty = bty;
  if (kc_out) *kc_out = bkc;
  if (row_out) *row_out = brow;
  if (ms_out) *ms_out = best;
  return 0;
}

}  // namespace

extern "C" {

int fuss_autotune(fuss_ctx* c, int level, const fuss_step_params* p, int steps, int* kc, int* row_march,
                  double* ms_per_step) {
  if (c) drop_dt_next(c);
  if (!c || !c->setup_done || !p || level < 1 || level > c->nlev || steps < 1) return FUSS_EARG;
  return autotune(c, level, p, steps, kc, row_march, ms_per_step);
}

int fuss_set_tiles(fuss_ctx* c, int level, int kc, int row_march, int stage_rows, int irs_rows) {
  if (c) drop_dt_next(c);
  if (!c || !c->setup_done || level < 1 || level > c->nlev) return FUSS_EARG;
  if ((kc != 4 && kc != 8 && kc != 16 && kc != 32 && kc != 64) || row_march < -1 || row_march > 1 ||
      (stage_rows != 4 && stage_rows != 8 && stage_rows != FUSS_STAGE_REGCOL) || (irs_rows != 4 && irs_rows != 8))
    return FUSS_EARG;
  Level& L = c->lev[level - 1];
  if (L.kp != GC) return 0;  // 2-D levels have no column tiles
  // the register-column kernel only where it can run (rectilinear level with
  // its per-axis factors in use, orthogonal blocks, tables from T >= 0, its
  // column heights)
  if (stage_rows == FUSS_STAGE_REGCOL &&
      !(L.rect_w && !c->rect_off && c->Ti1 >= 0 && !L.tang && L.dims2 && row_march != 1 &&
        (kc == 8 || kc == 16 || kc == 32)))
    return FUSS_EUNSUPPORTED;
  const int e = set_kc(c, L, kc);
  if (e) return e;
  L.row_march = row_march;
  L.stage_ty = stage_rows;
  L.xyz_ty = irs_rows;
  return 0;
}

int fuss_tiles(fuss_ctx* c, int level, int* kc, int* row_march, int* stage_rows, int* irs_rows) {
  if (!c || !c->setup_done || level < 1 || level > c->nlev) return FUSS_EARG;
  const Level& L = c->lev[level - 1];
  const bool col = L.kp == GC;
  if (kc) *kc = col ? L.kc : 0;
  if (row_march) *row_march = col ? L.row_march : 0;
  if (stage_rows) *stage_rows = col ? L.stage_ty : 0;
  if (irs_rows) *irs_rows = col ? L.xyz_ty : 0;
  return 0;
}

}  // extern "C"

namespace {

// Levels the persistent path runs: no IRS, no connection or chimera
// entries, no profiling, hooks or adjoint recording, at most 16384 tiles.
bool persistent_eligible(fuss_ctx* c, Level& L, const fuss_step_params* q) {
  if (q->irs_beta > 0.0 || c->prof.enabled || c->adj_store_on || L.nchim) return false;
  return L.htiles.size() <= 16384 && !L.has_conn;
}

// fuss_run through cooperative launches of up to PK_CHUNK steps each
// (persist.hip): the same steps, stop tests, history and final state as the
// step loop.
constexpr int PK_CHUNK = 1 << 15;

template <typename R>
int run_persistent(fuss_ctx* c, int level, const fuss_step_params* q, int maxcount, double thr, double tthr,
                   int* n_done, double* hist_res, double* hist_time, fuss_step_result* out) {
  Level& L = c->lev[level - 1];
  const int nrk = q->n_rk;
  if (nrk < 1 || nrk > 4) return FUSS_EARG;
  const int ntile = (int)L.htiles.size();
  const bool prim = q->integration == FUSS_PRIM;
  int grid = 0;
  int e = persistent_grid<R>(c->ndir, L.tang != 0, prim, ntile, &grid);
  if (e) return e;
  // control words (bar[2], ctl[8]), 2 x ntile minima and partials, the
  // history of one launch
  const bool hist = hist_res || hist_time;
  const size_t hsteps = hist ? (size_t)std::min(maxcount, PK_CHUNK) : 0;
  const size_t need = 64 + 4 * (size_t)ntile * sizeof(double) + 2 * hsteps * sizeof(double);
  if (L.pk.n < need && L.pk.alloc(need)) return FUSS_EHIP;
  unsigned char* base = L.pk.p;
  double* pmin2 = reinterpret_cast<double*>(base + 64);
  double* part2 = pmin2 + 2 * (size_t)ntile;
  double* hbuf = part2 + 2 * (size_t)ntile;
  const FieldView<R> V = fields<R>(c, L);
  // the first step's stage-1 input needs its ghosts (later states get
  // theirs from the stage that writes them)
  const bool ghosts = L.tang || L.ghost_read;
  if (ghosts && (e = launch_ghost(c, L, V.Tb[L.cur]))) return e;
  if (read_clock(c, L)) return FUSS_EHIP;
  PArgs<R> a{};
  StageArgs<R>& sa = a.sa;
  sa.blk = L.blk.p; sa.tiles = L.tiles.p; sa.dt = V.dt; sa.dtg = c->scal.p; sa.use_dtg = q->time_accurate;
  sa.qvol = V.qvol; sa.vol = V.vol; sa.mat = L.mat.p;
  for (int w = 0; w < 9; w++) sa.W[w] = V.W[w];
  sa.bc = L.bc.p; sa.fmap = L.fmap.p; sa.bcv = V.bcv; sa.bcf = V.bcf; sa.tab = dtab<R>(c);
  sa.ser = serdev(c); sa.clk = L.clk.p; sa.halt = nullptr; sa.status = c->status.p;
  DtArgs<R>& da = a.da;
  da.blk = L.blk.p; da.tiles = L.tiles.p; da.mat = L.mat.p; da.invL2 = V.invL2; da.dt = V.dt;
  da.tab = dtab<R>(c); da.vnn = q->vnn; da.cnt = L.cnt.p; da.rampa = q->rampa_iter; da.status = c->status.p;
  a.hot = L.bch.p;
  a.gmap = L.gmap.p;
  for (int b = 0; b < 3; b++) a.Tb[b] = V.Tb[b];
  a.nrk = nrk;
  a.jam = nrk == 4;
  for (int s = 0; s < nrk; s++) a.rkc[s] = nrk == 4 ? RK4A[s] : RKC[nrk - 1][s];
  a.ntile = ntile;
  a.time_accurate = q->time_accurate;
  a.ghosts = ghosts;
  a.compute_residual = 1;
  a.thr = thr;
  a.tthr = tthr;
  a.clk = L.clk.p;
  a.cnt = L.cnt.p;
  a.scal = c->scal.p;
  a.hres = hist ? hbuf : nullptr;
  a.htime = hist ? hbuf + hsteps : nullptr;
  a.pmin2 = pmin2;
  a.part2 = part2;
  a.bar = reinterpret_cast<unsigned*>(base);
  a.ctl = reinterpret_cast<int*>(base + 16);
  int total = 0;
  std::vector<double> hb(2 * hsteps);
  while (total < maxcount) {
    a.maxcount = std::min(maxcount - total, PK_CHUNK);
    a.cur0 = L.cur;
    a.time0 = L.time;
    a.iter0 = L.iter;
    HIPCHK(hipMemsetAsync(base, 0, 64, c->stream));
    e = launch_persistent<R>(c, a, c->ndir, L.tang != 0, prim, grid);
    if (e) return e;
    int ctl[8];
    HIPCHK(hipMemcpyAsync(ctl, a.ctl, sizeof(ctl), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    if (ctl[0]) {
      fprintf(stderr, "fuss_gpu: a grid barrier of the persistent launch timed out\n");
      return FUSS_EHIP;
    }
    const int done = ctl[2];
    if (hist && done > 0) {
      HIPCHK(hipMemcpy(hb.data(), hbuf, 2 * hsteps * sizeof(double), hipMemcpyDeviceToHost));
      for (int s = 0; s < done; s++) {
        if (hist_res) hist_res[total + s] = hb[s];
        if (hist_time) hist_time[total + s] = hb[hsteps + s];
      }
    }
    L.cur = ctl[3];
    total += done;
    if (read_clock(c, L)) return FUSS_EHIP;
    if (ctl[4] || done < a.maxcount) break;
  }
  L.dt_next = 0;
  int bijk[4] = {0, 0, 0, 0};
  HIPCHK(hipMemcpyAsync(c->hscal, c->scal.p, 3 * sizeof(double), hipMemcpyDeviceToHost, c->stream));
  const int st = sync_status(c, bijk);
  if (out) {
    memset(out, 0, sizeof(*out));
    out->res_sq = c->hscal[2];
    out->dtglobal = c->hscal[0];
    out->time = L.time;
    out->iter = L.iter;
    out->status = st;
    for (int k = 0; k < 4; k++) out->bijk[k] = bijk[k];
  }
  if (n_done) *n_done = total;
  return st;
}

}  // namespace

extern "C" {

static int step_sync(fuss_ctx* c, int level, const fuss_step_params* p, fuss_step_result* r, const StageHook* hk) {
  if (!c || !c->setup_done || !p || level < 1 || level > c->nlev) return FUSS_EARG;
  int e = enqueue_step(c, level, p, hk);
  if (e) return e;
  Level& L = c->lev[level - 1];
  HIPCHK(hipMemcpyAsync(c->hscal, c->scal.p, 3 * sizeof(double), hipMemcpyDeviceToHost, c->stream));
  HIPCHK(hipMemcpyAsync(c->hscal + 6, L.clk.p, sizeof(double), hipMemcpyDeviceToHost, c->stream));
  HIPCHK(hipMemcpyAsync(c->hcnt, L.cnt.p, 4 * sizeof(int), hipMemcpyDeviceToHost, c->stream));
  int bijk[4] = {0, 0, 0, 0};
  const int st = sync_status(c, bijk);
  L.time = c->hscal[6];
  L.iter = c->hcnt[0];
  if (r) {
    r->res_sq = p->compute_residual ? c->hscal[2] : 0.0;
    r->dtglobal = c->hscal[0];
    r->time = L.time;
    r->iter = L.iter;
    r->status = st;
    for (int q = 0; q < 4; q++) r->bijk[q] = bijk[q];
  }
  return st;
}

int fuss_step(fuss_ctx* c, int level, const fuss_step_params* p, fuss_step_result* r) {
  return step_sync(c, level, p, r, nullptr);
}

int fuss_step_hook(fuss_ctx* c, int level, const fuss_step_params* p, fuss_stage_fn fn, void* user,
                   fuss_step_result* r) {
  if (!c || !fn || c->adj_store_on) return FUSS_EARG;
  const StageHook hk{fn, user};
  return step_sync(c, level, p, r, &hk);
}

int fuss_run(fuss_ctx* c, int level, const fuss_step_params* p, int max_iters, int itermax_left,
             double res_threshold, double time_threshold, int* n_done, double* hist_res, double* hist_time,
             fuss_step_result* last) {
  if (!c || !c->setup_done || !p || level < 1 || level > c->nlev) return FUSS_EARG;
  fuss_step_result r;
  memset(&r, 0, sizeof(r));
  int n = 0;
  fuss_step_params q = *p;
  q.compute_residual = 1;
  // steps this call may take: the loop below stops after the first step with
  // n >= itermax_left, so at least one step runs when max_iters > 0
  const int maxcount = max_iters <= 0 ? 0 : std::min(max_iters, std::max(1, itermax_left));
  // captured chunk: RUN_G steps, or the request rounded up to a multiple of 3
  // (the steps past the limit are no-ops on the device); a few steps run
  // the loop
  const int G = maxcount >= RUN_G ? RUN_G : 3 * ((maxcount + 2) / 3);
  if (c->persist && maxcount >= 1 && persistent_eligible(c, c->lev[level - 1], &q)) {
    // one cooperative launch for all the steps (persist.hip)
    Level& L = c->lev[level - 1];
    if (c->precision == FUSS_FP32) {
      if (state_to_f32(c, L)) return FUSS_EHIP;
      return run_persistent<float>(c, level, &q, maxcount, res_threshold, time_threshold, n_done, hist_res,
                                   hist_time, last);
    }
    if (state_to_f64(c, L)) return FUSS_EHIP;
    return run_persistent<double>(c, level, &q, maxcount, res_threshold, time_threshold, n_done, hist_res,
                                  hist_time, last);
  }
  if (c->prof.enabled || maxcount < 6 || c->graphs_off) {
    while (n < max_iters) {
      const int e = fuss_step(c, level, &q, &r);
      if (e) {
        if (n_done) *n_done = n + 1;
        if (last) *last = r;
        return e;
      }
      const double res = sqrt(r.res_sq);
      if (hist_res) hist_res[n] = res;
      if (hist_time) hist_time[n] = r.time;
      n++;
      if (n >= itermax_left || res <= res_threshold || r.time >= time_threshold) break;
    }
    if (n_done) *n_done = n;
    if (last) *last = r;
    return 0;
  }
  // Graph mode: G steps (a multiple of 3, so the rotating T buffers come
  // back to their starting roles) captured once and replayed; the stop tests
  // run on the device (k_step_end) and later steps of the chunk become no-ops.
  Level& L = c->lev[level - 1];
  // the state in the precision of the captured steps (conversions stay
  // outside the capture)
  if (c->precision == FUSS_FP32) {
    if (L.nchim) return FUSS_EUNSUPPORTED;
    if (state_to_f32(c, L)) return FUSS_EHIP;
  } else if (state_to_f64(c, L)) {
    return FUSS_EHIP;
  }
  k_run_begin<<<1, 64, 0, c->stream>>>(L.cnt.p, maxcount);
  HIPCHK(hipGetLastError());
  if (q.irs_beta > 0.0 && q.irs_mode == FUSS_IRS_TRIDIAG) {
    const int e0 = tri_prepare(c, L, q.irs_beta);  // uploads: not inside the capture
    if (e0) return e0;
  }
  const bool hit = L.run_exec && !memcmp(&L.run_p, &q, sizeof(q)) && L.run_thr == res_threshold &&
                   L.run_tthr == time_threshold && L.run_g == G && L.run_bufs == run_buffers(c, L);
  if (hit && L.run_seq[0] != L.cur) {
    // the kept graph starts from another of the three rotating buffers: move
    // the state (T^n with its ghosts; the stage buffers are scratch) there
    const size_t bytes = L.npad * (c->precision == FUSS_FP32 ? sizeof(float) : sizeof(double));
    void* dst = c->precision == FUSS_FP32 ? (void*)L.f.Tb[L.run_seq[0]].p : (void*)L.Tb[L.run_seq[0]].p;
    const void* src = c->precision == FUSS_FP32 ? (const void*)L.f.Tb[L.cur].p : (const void*)L.Tb[L.cur].p;
    HIPCHK(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToDevice, c->stream));
    L.cur = L.run_seq[0];
  }
  int* cur_seq = L.run_seq;
  if (!hit) {
    run_graph_free(L);
    // the chunk's first step computes its dt: a replay may follow any change
    // of the state
    L.dt_next = 0;
    hipGraph_t graph = nullptr;
    HIPCHK(hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
    int e = 0;
    {
      k_chunk_begin<<<1, 64, 0, c->stream>>>(L.cnt.p);
      cur_seq[0] = L.cur;
      for (int s = 0; s < G && !e; s++) {
        e = enqueue_step(c, level, &q);
        k_step_end<<<1, 64, 0, c->stream>>>(c->scal.p, L.clk.p, L.cnt.p, c->status.p, c->run_hist.p,
                                            c->run_hist.p + RUN_G, res_threshold, time_threshold);
        cur_seq[s + 1] = L.cur;
      }
    }
    const hipError_t ce = hipStreamEndCapture(c->stream, &graph);
    L.cur = cur_seq[0];
    if (e || ce != hipSuccess) {
      if (graph) (void)hipGraphDestroy(graph);
      return e ? e : FUSS_EHIP;
    }
    if (hipGraphInstantiate(&L.run_exec, graph, nullptr, nullptr, 0) != hipSuccess) {
      (void)hipGraphDestroy(graph);
      L.run_exec = nullptr;
      return FUSS_EHIP;
    }
    L.run_graph = graph;
    L.run_bufs = run_buffers(c, L);
    L.run_g = G;
    L.run_p = q;
    L.run_thr = res_threshold;
    L.run_tthr = time_threshold;
  }
  hipGraphExec_t exec = L.run_exec;
  int e = 0;
  int done = 0, halted = 0;
  while (!halted) {
    if (hipGraphLaunch(exec, c->stream) != hipSuccess) { e = FUSS_EHIP; break; }
    HIPCHK(hipMemcpyAsync(c->hcnt, L.cnt.p, 4 * sizeof(int), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipMemcpyAsync(c->hrun, c->run_hist.p, 2 * RUN_G * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    const int k = c->hcnt[2] - done;
    for (int s = 0; s < k; s++) {
      if (hist_res) hist_res[done + s] = c->hrun[s];
      if (hist_time) hist_time[done + s] = c->hrun[RUN_G + s];
    }
    done += k;
    halted = c->hcnt[1];
    L.cur = cur_seq[halted ? k : G];
  }
  if (e) return e;
  HIPCHK(hipMemcpyAsync(c->hscal, c->scal.p, 3 * sizeof(double), hipMemcpyDeviceToHost, c->stream));
  HIPCHK(hipMemcpyAsync(c->hcnt, L.cnt.p, 4 * sizeof(int), hipMemcpyDeviceToHost, c->stream));
  HIPCHK(hipMemcpyAsync(c->hscal + 6, L.clk.p, sizeof(double), hipMemcpyDeviceToHost, c->stream));
  int bijk[4] = {0, 0, 0, 0};
  const int st = sync_status(c, bijk);
  // leave the halt flag down for the next caller
  k_run_begin<<<1, 64, 0, c->stream>>>(L.cnt.p, maxcount);
  L.time = c->hscal[6];
  L.iter = c->hcnt[0];
  r.res_sq = c->hscal[2];
  r.dtglobal = c->hscal[0];
  r.time = L.time;
  r.iter = L.iter;
  r.status = st;
  for (int q2 = 0; q2 < 4; q2++) r.bijk[q2] = bijk[q2];
  if (n_done) *n_done = done;
  if (last) *last = r;
  return st;
}

int fuss_restrict(fuss_ctx* c, int fine_level) {
  if (c) drop_dt_next(c);
  if (!c || !c->setup_done || fine_level < 1 || fine_level >= c->nlev) return FUSS_EARG;
  if (state_to_f64(c, c->lev[fine_level - 1])) return FUSS_EHIP;
  c->lev[fine_level].state_f = 0;  // overwritten in FP64
  int e = restrict_level(c, fine_level - 1);
  if (e) return e;
  return sync_status(c);
}

int fuss_prolong(fuss_ctx* c, int fine_level) {
  if (c) drop_dt_next(c);
  if (!c || !c->setup_done || fine_level < 1 || fine_level >= c->nlev) return FUSS_EARG;
  Level& F = c->lev[fine_level - 1];
  Level& C = c->lev[fine_level];
  if (state_to_f64(c, C)) return FUSS_EHIP;
  F.state_f = 0;  // overwritten in FP64
  for (int b = 0; b < F.nb; b++) {
    MGArgs a{};
    a.F = F.hblk[b]; a.Cb = C.hblk[b]; a.fT = F.Tb[F.cur].p; a.cT = C.Tb[C.cur].p;
    a.tab = dtab<double>(c); a.status = c->status.p; a.bidx = b;
    const long n = (long)a.Cb.ni * a.Cb.nj * a.Cb.nk;
    k_prolong<<<(unsigned)((n + 255) / 256), 256, 0, c->stream>>>(a);
    HIPCHK(hipGetLastError());
  }
  HIPCHK(hipStreamSynchronize(c->stream));
  return 0;
}

static int field_ptr(fuss_ctx* c, Level& L, const char* name, double** p) {
  // after FP32 steps the FP64 arrays of R and dt are refreshed from the
  // FP32 ones (the FP64 arrays are otherwise unused scratch then)
  const bool f = L.state_f != 0;
  if (!strcmp(name, "T")) {
    if (state_to_f64(c, L)) return FUSS_EHIP;
    *p = L.Tb[L.cur].p;
  } else if (!strcmp(name, "R")) {
    if (f && convert_f2d(c, L.f.R.p, L.R.p, L.npad)) return FUSS_EHIP;
    *p = L.R.p;
  } else if (!strcmp(name, "dt")) {
    if (f && convert_f2d(c, L.f.dt.p, L.dt.p, L.npad)) return FUSS_EHIP;
    *p = L.dt.p;
  } else if (!strcmp(name, "vol")) *p = L.vol.p;
  else if (!strcmp(name, "qvol")) *p = L.qvol.p;
  else return FUSS_EARG;
  return 0;
}

int fuss_get_field(fuss_ctx* c, int level, int b, const char* name, double* out) {
  if (!c || !c->setup_done || level < 1 || level > c->nlev || b < 1 || b > c->nb || !out || !name) return FUSS_EARG;
  Level& L = c->lev[level - 1];
  double* src = nullptr;
  const int ef = field_ptr(c, L, name, &src);
  if (ef) return ef;
  const BlockInfo& B = L.hblk[b - 1];
  std::vector<double> pad(B.npad);
  HIPCHK(hipMemcpyAsync(pad.data(), src + B.off, B.npad * 8, hipMemcpyDeviceToHost, c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  for (int k = 1; k <= B.nk; k++)
    for (int j = 1; j <= B.nj; j++)
      for (int i = 1; i <= B.ni; i++)
        out[(i - 1) + (size_t)B.ni * ((j - 1) + (size_t)B.nj * (k - 1))] = pad[cidx(B, i, j, k) - B.off];
  return 0;
}

int fuss_get_T(fuss_ctx* c, int level, int b, double* T) { return fuss_get_field(c, level, b, "T", T); }

int fuss_set_T(fuss_ctx* c, int level, int b, const double* T) {
  if (c) drop_dt_next(c);
  if (!c || !c->setup_done || level < 1 || level > c->nlev || b < 1 || b > c->nb || !T) return FUSS_EARG;
  Level& L = c->lev[level - 1];
  if (state_to_f64(c, L)) return FUSS_EHIP;
  const BlockInfo& B = L.hblk[b - 1];
  std::vector<double> pad(B.npad);
  HIPCHK(hipMemcpyAsync(pad.data(), L.Tb[L.cur].p + B.off, B.npad * 8, hipMemcpyDeviceToHost, c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  for (int k = 1; k <= B.nk; k++)
    for (int j = 1; j <= B.nj; j++)
      for (int i = 1; i <= B.ni; i++)
        pad[cidx(B, i, j, k) - B.off] = T[(i - 1) + (size_t)B.ni * ((j - 1) + (size_t)B.nj * (k - 1))];
  HIPCHK(hipMemcpy(L.Tb[L.cur].p + B.off, pad.data(), B.npad * 8, hipMemcpyHostToDevice));
  return launch_ghost(c, L, L.Tb[L.cur].p);
}

int fuss_probe(fuss_ctx* c, int n, const int* bijk, double* out) {
  if (!c || !c->setup_done || n < 0 || (n && (!bijk || !out))) return FUSS_EARG;
  Level& L = c->lev[0];
  if (state_to_f64(c, L)) return FUSS_EHIP;
  for (int q = 0; q < n; q++) {
    const int b = bijk[4 * q] - 1;
    if (b < 0 || b >= L.nb) return FUSS_EARG;
    const BlockInfo& B = L.hblk[b];
    const int i = bijk[4 * q + 1], j = bijk[4 * q + 2], k = bijk[4 * q + 3];
    if (i < 1 || i > B.ni || j < 1 || j > B.nj || k < 1 || k > B.nk) return FUSS_EARG;
    HIPCHK(hipMemcpyAsync(out + q, L.Tb[L.cur].p + cidx(B, i, j, k), 8, hipMemcpyDeviceToHost, c->stream));
  }
  HIPCHK(hipStreamSynchronize(c->stream));
  return 0;
}

int fuss_get_clock(fuss_ctx* c, int level, double* time, int* iter) {
  if (!c || level < 1 || level > c->nlev) return FUSS_EARG;
  if (c->setup_done && read_clock(c, c->lev[level - 1])) return FUSS_EHIP;
  if (time) *time = c->lev[level - 1].time;
  if (iter) *iter = c->lev[level - 1].iter;
  return 0;
}

int fuss_set_clock(fuss_ctx* c, int level, double time, int iter) {
  if (c) drop_dt_next(c);
  if (!c || level < 1 || level > c->nlev) return FUSS_EARG;
  c->lev[level - 1].time = time;
  c->lev[level - 1].iter = iter;
  if (c->setup_done) return set_clock(c, c->lev[level - 1], time, iter);
  return 0;
}

int fuss_ndir(fuss_ctx* c, int* ndir) {
  if (!c || !ndir || !c->setup_done) return FUSS_EARG;
  *ndir = c->ndir;
  return 0;
}

int fuss_get_metrics(fuss_ctx* c, int level, int b, double* M, double* vol, int d, double* N, double* A) {
  if (!c || !c->setup_done || level < 1 || level > c->nlev || b < 1 || b > c->nb) return FUSS_EARG;
  Level& L = c->lev[level - 1];
  const BlockInfo& B = L.hblk[b - 1];
  if (M || vol) {
    std::vector<double> m(9 * B.npad), v(B.npad);
    HIPCHK(hipMemcpy(m.data(), L.M.p + 9 * B.off, 9 * B.npad * 8, hipMemcpyDeviceToHost));
    HIPCHK(hipMemcpy(v.data(), L.vol.p + B.off, B.npad * 8, hipMemcpyDeviceToHost));
    for (int k = 1; k <= B.nk; k++)
      for (int j = 1; j <= B.nj; j++)
        for (int i = 1; i <= B.ni; i++) {
          const size_t s = (i - 1) + (size_t)B.ni * ((j - 1) + (size_t)B.nj * (k - 1));
          const long p = cidx(B, i, j, k) - B.off;
          if (M) for (int q = 0; q < 9; q++) M[9 * s + q] = m[9 * p + q];
          if (vol) vol[s] = v[p];
        }
  }
  if (d >= 1 && d <= 3 && (N || A)) {
    const int D = d - 1;
    std::vector<double> f(4 * B.npad);
    HIPCHK(hipMemcpy(f.data(), L.F[D].p + 4 * B.off, 4 * B.npad * 8, hipMemcpyDeviceToHost));
    const int e0 = B.ni + (D == 0), e1 = B.nj + (D == 1), e2 = B.nk + (D == 2);
    for (int kk = 0; kk < e2; kk++)
      for (int jj = 0; jj < e1; jj++)
        for (int ii = 0; ii < e0; ii++) {
          const int i = D == 0 ? ii : ii + 1, j = D == 1 ? jj : jj + 1, k = D == 2 ? kk : kk + 1;
          const long p = cidx(B, i, j, k) - B.off;
          const size_t s = ii + (size_t)e0 * (jj + (size_t)e1 * kk);
          if (N) for (int q = 0; q < 3; q++) N[3 * s + q] = f[4 * p + q];
          if (A) A[s] = f[4 * p + 3];
        }
  }
  return 0;
}

int fuss_get_stream(fuss_ctx* c, void** s) {
  if (!c || !s) return FUSS_EARG;
  *s = (void*)c->stream;
  return 0;
}

int fuss_set_stream(fuss_ctx* c, void* s) {
  if (!c) return FUSS_EARG;
  if (c->own_stream && c->stream) (void)hipStreamDestroy(c->stream);
  c->stream = (hipStream_t)s;
  c->own_stream = false;
  return 0;
}

int fuss_profile(fuss_ctx* c, int enable) {
  if (!c) return FUSS_EARG;
  c->prof.enabled = enable != 0;
  c->prof.used = 0;
  c->prof.cls.clear();
  return 0;
}

int fuss_profile_read(fuss_ctx* c, int nclass, double* ms, int* count) {
  if (!c || nclass < 1 || !ms || !count) return FUSS_EARG;
  HIPCHK(hipStreamSynchronize(c->stream));
  for (int q = 0; q < nclass; q++) {
    ms[q] = 0.0;
    count[q] = 0;
  }
  auto& P = c->prof;
  for (size_t id = 0; id < P.used; id++) {
    float t = 0.f;
    HIPCHK(hipEventElapsedTime(&t, P.pool[2 * id], P.pool[2 * id + 1]));
    if (P.cls[id] < nclass) {
      ms[P.cls[id]] += t;
      count[P.cls[id]] += 1;
    }
  }
  P.used = 0;
  P.cls.clear();
  return 0;
}

int fuss_memory(fuss_ctx* c, double* bytes) {
  if (!c || !bytes) return FUSS_EARG;
  double s = 0;
  for (auto& L : c->lev) {
    s += 8.0 * (3 + 6) * L.npad + 4.0 * L.npad + 8.0 * (9 + 12 + 9) * L.npad + 24.0 * L.nnodes;
    s += sizeof(BCRec) * (double)L.nbound;
  }
  *bytes = s;
  return 0;
}

void fuss_destroy(fuss_ctx* c) {
  if (!c) return;
  if (c->stream) (void)hipStreamSynchronize(c->stream);
  for (auto& L : c->lev) {
    run_graph_free(L);
    L.ttx.free(); L.tty.free(); L.ttz.free(); L.tri.free(); L.trioff.free();
    L.blk.free(); L.tiles.free(); L.nodes.free(); L.rfac.free();
    for (auto& t : L.Tb) t.free();
    L.R.free(); L.R2.free(); L.dt.free(); L.qvol.free(); L.vol.free(); L.invL2.free(); L.mat.free();
    L.M.free();
    for (auto& f : L.F) f.free();
    for (auto& w : L.W) w.free();
    L.bc.free(); L.bcv.free(); L.bcf.free(); L.bch.free(); L.gmap.free(); L.fmap.free(); L.xytiles.free(); L.partials.free(); L.pmin.free(); L.red2.free(); L.ctiles.free();
    L.c8tiles.free();
    for (auto& r : L.rng) r.free();
    L.chr.free(); L.chd.free(); L.chv.free();
  }
  c->dcp.free(); c->drho.free(); c->dkap.free(); c->dh.free(); c->dkap2.free(); c->dpr4.free();
  c->status.free(); c->scal.free();
  refine_free(c->ref);
  adj_free(c->adj);
  c->stage_i.free(); c->stage_d.free();
  c->ser_t.free(); c->ser_v.free(); c->ser_off.free(); c->ser_n.free(); c->ser_per.free(); c->run_hist.free();
  for (auto& L : c->lev) {
    L.clk.free(); L.cnt.free();
    for (auto& r : L.rng) r.free();
  }
  if (c->hrun) (void)hipHostFree(c->hrun);
  if (c->hcnt) (void)hipHostFree(c->hcnt);
  if (c->hscal) (void)hipHostFree(c->hscal);
  if (c->hstatus) (void)hipHostFree(c->hstatus);
  for (auto e : c->prof.pool) (void)hipEventDestroy(e);
  if (c->own_stream && c->stream) (void)hipStreamDestroy(c->stream);
  delete c;
}

}  // extern "C"
