// fuss_dev.h -- argument blocks and device helpers shared by the stage,
// IRS and surface kernels of the FUSS explicit step (gfx950), and the
// HIPCHK status macro of the host code.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdio>
#include <cstring>

#include "../../include/fuss_gpu.h"
#include "fuss_kernels.h"

using namespace fuss;

#define HIPCHK(x)                                                     \
  do {                                                                \
    hipError_t e_ = (x);                                              \
    if (e_ != hipSuccess) {                                           \
      fprintf(stderr, "fuss_gpu: HIP error %s at %s:%d\n",            \
              hipGetErrorString(e_), __FILE__, __LINE__);             \
      return FUSS_EHIP;                                               \
    }                                                                 \
  } while (0)


// ------------------------------------------------------------------------
// kernels: local dt (per iteration)
// ------------------------------------------------------------------------
template <typename R>
struct DtArgs {
  const BlockInfo* blk;
  const Tile* tiles;
  const R* T;
  const int* mat;
  const double* invL2;
  R* dt;
  Tab<R> tab;
  double vnn;
  const int* cnt;      // level counters: [0] iteration before this step, [1] halt
  int rampa;
  double* pmin;        // per-work-group minimum
  int* status;
};

// ------------------------------------------------------------------------
// kernels: stage (fluxes + BC fluxes + newstate)
// ------------------------------------------------------------------------
template <typename R>
struct StageArgs {
  const BlockInfo* blk;
  const Tile* tiles;
  const Tile* ctiles;  // column tiles (k0 = first plane of a KC-plane chunk)
  const R* T;          // stage input (ghosts filled)
  const R* TO;         // T^n
  R* Tout;             // stage output (update mode)
  R* Rres;             // scaled residual (IRS mode)
  const R* dt;
  const double* dtg;   // device dtglobal (time-accurate)
  int use_dtg;
  const R* qvol;
  const R* vol;
  const int* mat;
  const R* W[9];
  const BCRec* bc;
  const int* fmap;
  R* bcv;              // per BC entry: value subtracted from the cell residual
  R* bcf;              // the same values at the entries' face-map positions
  Tab<R> tab;
  SerDev ser;          // 301/302 time series, evaluated at *clk
  const double* clk;   // level time (device)
  const int* halt;     // nonzero: the step is skipped (graph replay past a stop)
  R rkc;
  int jam;             // four-stage Jameson update (n_rk = 4)
  int last;
  int irk;             // stage number (1-based) of the step, for the state-check location
  double* partials;
  int* status;
  const double* rect;  // per-axis metric factors (rectilinear levels, k_stage3p<RECT>)
  // k_stage3p<DTC>: the first stage of a steady step computes the local dt
  // from its input (T^n) and stores it for the later stages
  double vnn;
  R* dtw;
  double* pmin;        // per-tile dt minima
  int conn_split;      // connection (101) entries are evaluated by k_connflux, not k_bcflux
  unsigned bcz;        // an entry of bcf that is always +0 (its last)
};

// Compute_dt (Mod_dt.f90:26-84) for the cells of tile `wg` at iteration
// `iter` (vnn ramp): local dt stored, the tile's minimum to pmin[wg].
template <typename R>
__device__ __forceinline__ void dt_tile(const DtArgs<R>& a, int wg, int iter) {
  const Tile t = a.tiles[wg];
  const BlockInfo B = a.blk[t.b];
  const int i = t.i0 + (threadIdx.x % TILE_X), j = t.j0 + (threadIdx.x / TILE_X), k = t.k0;
  double dtc = 1e300;
  if (i <= B.ni && j <= B.nj) {
    const long p = B.base + (i - 1) + (long)(j - 1) * B.sy + (long)(k - 1) * B.sz;
    const R T = a.T[p];
    const int m = B.umat >= 0 ? B.umat : a.mat[p];
    int bad = 0;
    const long row = (long)m * a.tab.nT;
    const R kap = tab_interp(a.tab.kap + row, T, a.tab.Ti1, a.tab.Ti2, &bad);
    const R rho = tab_interp(a.tab.rho + row, T, a.tab.Ti1, a.tab.Ti2, &bad);
    const R cp = tab_interp(a.tab.cp + row, T, a.tab.Ti1, a.tab.Ti2, &bad);
    {
#pragma clang fp contract(off)
      const double alpha = (double)kap / (double)rho / (double)cp;
      double d = 0.5 * a.vnn / (alpha * a.invL2[p]);
      d = d < 1e8 ? d : 1e8;
      if (iter < a.rampa) d = d * (double)iter / (double)a.rampa;
      dtc = d;
    }
    a.dt[p] = (R)dtc;
    if (bad) flag_error(a.status, FUSS_ETABLE, t.b, i, j, k);
  }
  // work-group minimum
  __shared__ double sm[TILE_T / 64];
  for (int o = 32; o > 0; o >>= 1) dtc = fmin(dtc, __shfl_xor(dtc, o, 64));
  if ((threadIdx.x & 63) == 0) sm[threadIdx.x >> 6] = dtc;
  __syncthreads();
  if (threadIdx.x == 0) {
    double v = sm[0];
    for (int q = 1; q < TILE_T / 64; q++) v = fmin(v, sm[q]);
    a.pmin[wg] = v;
  }
  __syncthreads();  // sm is reused by the next call in a persistent loop
}

// One interior face flux between cells q and q+e_D (left cell q), in the
// gather form of Diffusive_Flux (Lib_Diffusive.f90:10-45).
template <int D, int NDIR, bool TANG, typename R>
__device__ __forceinline__ R face_flux(const StageArgs<R>& a, const BlockInfo& B, long q) {
  const R* __restrict__ T = a.T;
  const long sy = B.sy, sz = B.sz;
  const long s = D == 0 ? 1 : (D == 1 ? sy : sz);
  const R Tl = T[q], Tr = T[q + s];
  R G0, G1 = 0, G2 = 0;
  G0 = Tr - Tl;
  if (TANG) {
    // tangential differences, Tangential_Gradient (Lib_Diffusive.f90:48-55)
    const long t1 = D == 0 ? sy : 1;
    const long t2 = D == 2 ? sy : sz;
    G1 = (T[q + t1] - T[q - t1] + T[q + s + t1] - T[q + s - t1]) * (R)0.25;
    if (NDIR == 3 || D == 2) G2 = (T[q + t2] - T[q - t2] + T[q + s + t2] - T[q + s - t2]) * (R)0.25;
  }
  int bad = 0;
  const R kap = tab_interp(a.tab.kap + (long)a.mat[q] * a.tab.nT, (R)0.5 * (Tl + Tr), a.tab.Ti1,
                           a.tab.Ti2, &bad);
  if (bad) flag_error(a.status, FUSS_ETABLE, -1, 0, 0, 0);
  // weights in computational-direction order; the normal direction is D
  const R wn = a.W[3 * D + D][q];
  if (!TANG) return kap * (G0 * wn);
  const int c1 = D == 0 ? 1 : 0, c2 = D == 2 ? 1 : 2;
  const R w1 = a.W[3 * D + c1][q], w2 = a.W[3 * D + c2][q];
  // sum over c = 0,1,2 in order of the computational index
  R g[3], w[3];
  g[D] = G0; w[D] = wn;
  g[c1] = G1; w[c1] = w1;
  g[c2] = G2; w[c2] = w2;
  return kap * ((g[0] * w[0] + g[1] * w[1]) + g[2] * w[2]);
}

template <typename R>
__device__ __forceinline__ R dotrow(const double* M, int r, const double* N) {
  // dot_product(M(r,:), N), r = 1..3
  return (R)(M[r - 1] * N[0] + M[r - 1 + 3] * N[1] + M[r - 1 + 6] * N[2]);
}

// Wall fluxes that need only scalars of their entry; bc_value (full record)
// and the compact-record pass of k_bcflux both evaluate them here.
// BC_Wall_HeatTransfer, Lib_BC_Fluxes_Wall_HeatTransfer.f90:12-52 (the
// reference passes bc%qw, never set for 303: 0 here).
template <typename R>
__device__ __forceinline__ R bcf_htc(R kappa, R mn, R A, R hconv, R Tref, R Tc) {
#pragma clang fp contract(off)
  const R H = (R)1.0 / ((R)2.0 * kappa * mn) + (R)1.0 / hconv;
  const R Hrad = hconv / ((R)2.0 * kappa * mn) + (R)1.0;
  return A * ((R)1.0 / H * (Tref - Tc) + (R)1.0 / Hrad * (R)0.0);
}

// BC_Wall_Radiation(+Convection), Lib_BC_Fluxes_Wall_Radiation.f90:12-65 and
// Lib_BC_Fluxes_Wall_RadiationConvection.f90:12-65.  The fixed point stays in
// FP64: its tolerance (1e-3 W/m^2) is below the FP32 resolution of q.
// Returns the flux; *fail is set when the iteration does not settle.
__device__ __forceinline__ double bcf_rad(int type, double kappa, double mn, double A, double Tref, double eps,
                                          double hconv, double Tcd, int* fail) {
#pragma clang fp contract(off)
  const double Tr2 = Tref * Tref, Tr4 = Tr2 * Tr2;
  double Tw = Tcd, qw = 0.0, err = 1.0e6, Flux = 0.0;
  int it = 0;
  while (err > 1.0e-3) {
    const double Tw2 = Tw * Tw;
    const double K = SIGMA_SB * eps * (Tw2 * Tw2 - Tr4) / (Tw - Tref);
    const double H = type == 304 ? 1.0 / (2.0 * kappa * mn) + 1.0 / K : 1.0 / (2.0 * kappa * mn) + 1.0 / (K + hconv);
    Flux = A * (1.0 / H * (Tref - Tcd));
    const double q_wall = K * (Tref - Tw);
    err = fabs(qw - q_wall);
    qw = q_wall;
    Tw = Tref - 1.0 / (H * K) * (Tref - Tcd);
    if (++it > 1000 || err != err) {
      *fail = 1;
      break;
    }
  }
  return Flux;
}

// BC_Connection (Lib_BC_Fluxes_Connection.f90:12-113) for one entry: the
// interface temperature from both sides' conductances, then the flux of the
// gradient through it, in the reference's order.  p: the boundary cell;
// gsrc: its layer-1 ghost's donor; dn: the donor cells of Tg(3:6); M, Nr: the
// cell's metric and the face normal; mc = Mg(1) . N; krow / kcon: the kappa
// rows of the cell's and the ghost's materials.  Returns v with R = R - v.
template <int NDIR, typename R>
__device__ __forceinline__ R conn_value(const R* __restrict__ T, long p, long sy, long sz, int Fm, const double* M,
                                        const double* Nr, const double* mc3, double A, long gsrc, const long* dn,
                                        const R* krow, const R* kcon, int Ti1, int Ti2, int* bad) {
#pragma clang fp contract(off)
  const int modfm2 = 1 - 2 * (Fm % 2);
  const R Tc = T[p];
  const R Tip = T[p + 1], Tim = T[p - 1], Tjp = T[p + sy], Tjm = T[p - sy];
  R Tkp = 0, Tkm = 0;
  if (NDIR == 3) { Tkp = T[p + sz]; Tkm = T[p - sz]; }
  const R Tcon = T[gsrc];
  const R kl = tab_interp(krow, Tc, Ti1, Ti2, bad);
  const R kc = tab_interp(kcon, Tcon, Ti1, Ti2, bad);
  const R g3 = T[dn[0]], g4 = T[dn[1]], g5 = T[dn[2]], g6 = T[dn[3]];
  const R ml0 = dotrow<R>(M, 1, Nr), ml1 = dotrow<R>(M, 2, Nr), ml2 = dotrow<R>(M, 3, Nr);
  const R mc0 = (R)mc3[0], mc1 = (R)mc3[1], mc2 = (R)mc3[2];
  R Dl, Nl, Dc, Nc;
  if (Fm <= 2) {
    Dl = (R)2.0 * ml0 * modfm2;
    Nl = (R)2.0 * Tc * ml0 * modfm2 - (R)0.5 * (Tjp - Tjm) * ml1 - (R)0.5 * (Tkp - Tkm) * ml2;
    Dc = (R)2.0 * mc0 * modfm2;
    Nc = (R)2.0 * Tcon * mc0 * modfm2 + (R)0.5 * (g4 - g3) * mc1 + (R)0.5 * (g6 - g5) * mc2;
  } else if (Fm <= 4) {
    Dl = (R)2.0 * ml1 * modfm2;
    Nl = -(R)0.5 * (Tip - Tim) * ml0 + (R)2.0 * Tc * ml1 * modfm2 - (R)0.5 * (Tkp - Tkm) * ml2;
    Dc = (R)2.0 * mc1 * modfm2;
    Nc = (R)0.5 * (g4 - g3) * mc0 + (R)2.0 * Tcon * mc1 * modfm2 + (R)0.5 * (g6 - g5) * mc2;
  } else {
    Dl = (R)2.0 * ml2 * modfm2;
    Nl = -(R)0.5 * (Tip - Tim) * ml0 - (R)0.5 * (Tjp - Tjm) * ml1 + (R)2.0 * Tc * ml2 * modfm2;
    Dc = (R)2.0 * mc2 * modfm2;
    Nc = (R)0.5 * (g4 - g3) * mc0 + (R)0.5 * (g6 - g5) * mc1 + (R)2.0 * Tcon * mc2 * modfm2;
  }
  const R Tint = (kl * Nl + kc * Nc) / (kl * Dl + kc * Dc);
  R G[3];
  if (Fm <= 2) { G[0] = (R)2.0 * (Tint - Tc) * modfm2; G[1] = (Tjp - Tjm) * (R)0.5; G[2] = (Tkp - Tkm) * (R)0.5; }
  else if (Fm <= 4) { G[0] = (Tip - Tim) * (R)0.5; G[1] = (R)2.0 * (Tint - Tc) * modfm2; G[2] = (Tkp - Tkm) * (R)0.5; }
  else { G[0] = (Tip - Tim) * (R)0.5; G[1] = (Tjp - Tjm) * (R)0.5; G[2] = (R)2.0 * (Tint - Tc) * modfm2; }
  if (NDIR == 2) G[2] = 0;
  R gr[3];
  for (int x = 0; x < 3; x++) gr[x] = G[0] * (R)M[3 * x] + G[1] * (R)M[3 * x + 1] + G[2] * (R)M[3 * x + 2];
  const R Flux = (R)A * kl * (gr[0] * (R)Nr[0] + gr[1] * (R)Nr[1] + gr[2] * (R)Nr[2]);
  return (R)modfm2 * Flux;
}

// BC_Fluxes for one entry (Mod_BC_Fluxes.f90:43-79 and the per-type files
// src/lib/numerics/fluxes/bc/Lib_BC_Fluxes_*.f90), in the reference's
// arithmetic order.
// Returns v such that the reference's residual update is R = R - v; the stage
// kernels apply the v of a cell's faces in face order, which reproduces the
// reference's accumulation exactly.
template <int NDIR, typename R>
__device__ R bc_value(const StageArgs<R>& a, const BlockInfo& B, const BCRec& e, long p, double sv) {
#pragma clang fp contract(off)
  R Rv = 0;
  const R* __restrict__ T = a.T;
  const int Fm = e.f;
  const int modfm2 = 1 - 2 * (Fm % 2);
  const int modfm3 = 2 * (2 * (Fm % 2) - 1);
  const R Tc = T[p];
  const long sy = B.sy, sz = B.sz;
  const double* M = e.M;
  const double* Nr = e.N;
  int bad = 0;
  const R* krow = a.tab.kap + (long)a.mat[p] * a.tab.nT;
  switch (e.type) {
    case 101:
    case 102:
      // BC_Connection, Lib_BC_Fluxes_Connection.f90:12-113 (chimera:
      // Mod_BC_Fluxes.f90:45, with its ghost as Tg(1) and the extrapolated Mg)
      Rv = conn_value<NDIR, R>(T, p, sy, sz, Fm, M, Nr, e.mc, e.A, e.gsrc, e.dn, krow,
                               a.tab.kap + (long)e.matg * a.tab.nT, a.tab.Ti1, a.tab.Ti2, &bad);
      break;
    case 300: {
      // BC_Symmetry, Lib_BC_Fluxes_Symmetry.f90:12-67
      R G[3] = {0, 0, 0};
      if (Fm > 2) G[0] = (R)0.5 * (T[p + 1] - T[p - 1]) * modfm2;
      if (Fm <= 2 || Fm > 4) G[1] = (R)0.5 * (T[p + sy] - T[p - sy]) * modfm2;
      if (NDIR == 3 && Fm <= 4) G[2] = (R)0.5 * (T[p + sz] - T[p - sz]) * modfm2;
      R gr[3];
      for (int x = 0; x < 3; x++) gr[x] = G[0] * (R)M[3 * x] + G[1] * (R)M[3 * x + 1] + G[2] * (R)M[3 * x + 2];
      const R kappa = tab_interp(krow, Tc, a.tab.Ti1, a.tab.Ti2, &bad);
      const R Flux = (R)e.A * kappa * (gr[0] * (R)Nr[0] + gr[1] * (R)Nr[1] + gr[2] * (R)Nr[2]);
      Rv = (R)modfm2 * Flux;
    } break;
    case 301: {
      // BC_Wall_Heat, Lib_BC_Fluxes_Wall_Heat.f90:12-54
      const R qw = e.series >= 0 ? (R)sv : (R)e.qw;
      const R Flux = (R)e.A * qw;
      Rv = (R)modfm2 * Flux;
    } break;
    case 302: {
      // BC_Wall_Temperature, Lib_BC_Fluxes_Wall_Temperature.f90:12-68
      const R Tw = e.series >= 0 ? (R)sv : (R)e.Tw;
      const R kappa = tab_interp(krow, Tc, a.tab.Ti1, a.tab.Ti2, &bad);
      R G[3] = {0, 0, 0};
      const R gn = (Tc - Tw) * (R)modfm3;
      if (e.dir == 1) G[0] = gn;
      else if (e.dir == 2) G[1] = gn;
      else G[2] = gn;
      if (Fm <= 2) {
        G[1] = (T[p - sy] - T[p + sy]) / (R)2.0;
        if (NDIR == 3) G[2] = (T[p - sz] - T[p + sz]) / (R)2.0;
      } else if (Fm <= 4) {
        G[0] = (T[p - 1] - T[p + 1]) / (R)2.0;
        if (NDIR == 3) G[2] = (T[p - sz] - T[p + sz]) / (R)2.0;
      } else {
        G[0] = (T[p - 1] - T[p + 1]) / (R)2.0;
        G[1] = (T[p - sy] - T[p + sy]) / (R)2.0;
      }
      if (NDIR == 2) G[2] = 0;
      R gr[3];
      for (int x = 0; x < 3; x++) gr[x] = G[0] * (R)M[3 * x] + G[1] * (R)M[3 * x + 1] + G[2] * (R)M[3 * x + 2];
      const R Flux = (R)e.A * kappa * (gr[0] * (R)Nr[0] + gr[1] * (R)Nr[1] + gr[2] * (R)Nr[2]);
      Rv = (R)modfm2 * Flux;
    } break;
    case 303: {
      const R kappa = tab_interp(krow, Tc, a.tab.Ti1, a.tab.Ti2, &bad);
      Rv = bcf_htc<R>(kappa, dotrow<R>(M, e.dir, Nr), (R)e.A, (R)e.hconv, (R)e.Tref, Tc);
    } break;
    case 304:
    case 305: {
      const double kappa = (double)tab_interp(krow, Tc, a.tab.Ti1, a.tab.Ti2, &bad);
      const double mn = M[e.dir - 1] * Nr[0] + M[e.dir - 1 + 3] * Nr[1] + M[e.dir - 1 + 6] * Nr[2];
      int fail = 0;
      Rv = (R)bcf_rad(e.type, kappa, mn, e.A, e.Tref, e.eps, e.hconv, (double)Tc, &fail);
      if (fail) flag_error(a.status, FUSS_ESTATE, e.b, e.i, e.j, e.k);
    } break;
    default:
      break;
  }
  if (bad) flag_error(a.status, FUSS_ETABLE, e.b, e.i, e.j, e.k);
  return Rv;
}

template <typename R>
__device__ __forceinline__ int face_map_index(const BlockInfo& B, int f, int i, int j, int k) {
  if (f <= 2) return B.fmap_off[f - 1] + (j - 1) + B.nj * (k - 1);
  if (f <= 4) return B.fmap_off[f - 1] + (i - 1) + B.ni * (k - 1);
  return B.fmap_off[f - 1] + (i - 1) + B.ni * (j - 1);
}

// BC entries of a boundary cell in the (block, face) order of the BC list
// (Mod_BC_Fluxes.f90:29-33): faces 1..6 in turn.
template <typename R>
__device__ __forceinline__ R apply_bc_faces(const StageArgs<R>& a, const BlockInfo& B, int i, int j, int k, R Rv) {
#pragma clang fp contract(off)
  const int fa = B.face_active;
  if (!fa) return Rv;
  if ((fa & 1) && i == 1) Rv = Rv - a.bcf[B.fmap_off[0] + (j - 1) + B.nj * (k - 1)];
  if ((fa & 2) && i == B.ni) Rv = Rv - a.bcf[B.fmap_off[1] + (j - 1) + B.nj * (k - 1)];
  if ((fa & 4) && j == 1) Rv = Rv - a.bcf[B.fmap_off[2] + (i - 1) + B.ni * (k - 1)];
  if ((fa & 8) && j == B.nj) Rv = Rv - a.bcf[B.fmap_off[3] + (i - 1) + B.ni * (k - 1)];
  if ((fa & 16) && k == 1) Rv = Rv - a.bcf[B.fmap_off[4] + (i - 1) + B.ni * (j - 1)];
  if ((fa & 32) && k == B.nk) Rv = Rv - a.bcf[B.fmap_off[5] + (i - 1) + B.ni * (j - 1)];
  return Rv;
}

// The value BC entry e subtracts from its cell's residual (BC_Fluxes,
// Mod_BC_Fluxes.f90:43-79, and the per-type files) at level time `time`:
// 301/303/304/305 from the compact record, 101/102/300/302 (tangential
// gradients, donor data) from the full record; 0 for entries without a flux.
template <int NDIR, typename R>
__device__ __forceinline__ R bc_entry_value(const StageArgs<R>& a, const BCHot* __restrict__ hot, int e,
                                            double time) {
  const int type = hot[e].type;
  if (!(type == 101 || type == 102 || (type >= 300 && type <= 305))) return (R)0;
  if (type == 301 || type >= 303) {
#pragma clang fp contract(off)
    const BCHot h = hot[e];
    const R Tc = ldu(a.T, h.cell);
    int bad = 0, fail = 0;
    R v;
    if (type == 301) {
      // BC_Wall_Heat, Lib_BC_Fluxes_Wall_Heat.f90:12-54 (signed along +xi)
      R qw = (R)h.qw;
      if (h.series >= 0) {
        const int q = h.series;
        qw = (R)series_value(a.ser.per[q], a.ser.n[q], a.ser.t + a.ser.off[q], a.ser.v + a.ser.off[q], time);
      }
      const R Flux = (R)h.A * qw;
      v = (R)(1 - 2 * (h.f % 2)) * Flux;
    } else {
      const R kappa = tab_interp(a.tab.kap + (long)h.m * a.tab.nT, Tc, a.tab.Ti1, a.tab.Ti2, &bad);
      if (type == 303) v = bcf_htc<R>(kappa, (R)h.mn, (R)h.A, (R)h.hconv, (R)h.Tref, Tc);
      else v = (R)bcf_rad(type, (double)kappa, h.mn, h.A, h.Tref, h.eps, h.hconv, (double)Tc, &fail);
    }
    if (bad || fail) {
      const BCRec& r = a.bc[e];
      flag_error(a.status, fail ? FUSS_ESTATE : FUSS_ETABLE, r.b, r.i, r.j, r.k);
    }
    return v;
  }
  const BCRec& r = a.bc[e];
  const BlockInfo B = a.blk[r.b];
  double sv = 0.0;
  if (r.series >= 0) {
    const int q = r.series;
    sv = series_value(a.ser.per[q], a.ser.n[q], a.ser.t + a.ser.off[q], a.ser.v + a.ser.off[q], time);
  }
  return bc_value<NDIR>(a, B, r, r.cell, sv);
}

// BC_Fluxes (Mod_BC_Fluxes.f90:10-86) as a surface pass: one thread per
// entry stores the value its face subtracts from the cell residual.
// Entries of type 301/303/304/305 are evaluated from their compact record;
// 101, 300 and 302 (tangential gradients, donor data) from the full record.
// RK_stage, Lib_RK.f90:15-26.  jam: the four-stage scheme of Jameson,
// Schmidt and Turkel (new = old + a_k res, a = 1/4, 1/3, 1/2, 1), which has
// no counterpart in the reference (its RKcoeff stops at three stages).
template <typename R>
__device__ __forceinline__ R rk_stage(R state, R old, R res, R c, int jam) {
#pragma clang fp contract(off)
  if (jam) return old + res * c;
  const R r_ = (state - old + res) * c;
  return old + r_;
}

// Newstate for one cell from its scaled residual.  PRIM: Newstate_Primitive
// (Lib_Newstate.f90:241-270); else Newstate_Conservative (:99-146).
template <bool PRIM, typename R>
__device__ __forceinline__ R new_state(const Tab<R>& tab, int m, R T, R TO, R res, R c, int jam, int* bad) {
#pragma clang fp contract(off)
  if (PRIM) return rk_stage(T, TO, res, c, jam);
  const long row = (long)m * tab.nT;
  const R hO = tab_interp(tab.h + row, TO, tab.Ti1, tab.Ti2, bad);
  R h = tab_interp(tab.h + row, T, tab.Ti1, tab.Ti2, bad);
  h = rk_stage(h, hO, res, c, jam);
  return tab_inverse(tab.h + row, h, T, tab.Ti1, tab.Ti2, bad);
}

template <bool PRIM, typename R>
__device__ __forceinline__ R scaled_residual(const Tab<R>& tab, int m, R T, R Rv, R vol, R qvol, R dt,
                                             int* bad) {
#pragma clang fp contract(off)
  if (!PRIM) return (-Rv / vol + qvol) * dt;  // Scale_Residual_Cons, Lib_Newstate.f90:99-104
  // co_Rho / co_Cp / co_DrhoDT / co_DcpDT from one (rho, drho, cp, dcp) row
  R Tc = T;
  if (!(Tc >= (R)tab.Ti1 && Tc < (R)tab.Ti2)) {
    *bad = 1;
    Tc = Tc < (R)tab.Ti1 || Tc != Tc ? (R)tab.Ti1 : (R)(tab.Ti2 - 1);
  }
  const int Tm = (int)Tc;
  const typename Vec4<R>::T v =
      ldu(reinterpret_cast<const typename Vec4<R>::T*>(tab.pr4), (unsigned)((long)m * tab.nT + (Tm - tab.Ti1)));
  const R rho = v.x + (Tc - (R)Tm) * v.y;
  const R cp = v.z + (Tc - (R)Tm) * v.w;
  const R drho = v.y, dcp = v.w;
  // Compute_Prim_Residual, Lib_Newstate.f90:241-256
  return (-Rv / vol + qvol) * dt / (T * cp * drho + rho * cp + T * rho * dcp);
}

// Work-group minimum to pmin[wg] (as k_dt's per-tile minimum).
// NT: threads of the work-group (a multiple of 64)
template <int NT = TILE_T>
__device__ __forceinline__ void block_min(double v, double* pmin, int wg) {
  __shared__ double sm[NT / 64];
  for (int o = 32; o > 0; o >>= 1) v = fmin(v, __shfl_xor(v, o, 64));
  if ((threadIdx.x & 63) == 0) sm[threadIdx.x >> 6] = v;
  __syncthreads();
  if (threadIdx.x == 0) {
    double s = sm[0];
    for (int q = 1; q < NT / 64; q++) s = fmin(s, sm[q]);
    pmin[wg] = s;
  }
}

template <int NT = TILE_T>
__device__ __forceinline__ void block_sum(double v, double* partials, int wg) {
  __shared__ double sm[NT / 64];
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  if ((threadIdx.x & 63) == 0) sm[threadIdx.x >> 6] = v;
  __syncthreads();
  if (threadIdx.x == 0) {
    double s = sm[0];
    for (int q = 1; q < NT / 64; q++) s += sm[q];
    partials[wg] = s;
  }
}

// IRSMODE: write the scaled residual to Rres (smoothing follows);
// otherwise apply the RK update to Tout.
constexpr int KC_DEFAULT = 8;   // column height; fuss_autotune may choose 4, 16, 32 or 64

// ------------------------------------------------------------------------
// kernels: implicit residual smoothing
// ------------------------------------------------------------------------
template <typename R>
struct IrsArgs {
  const BlockInfo* blk;
  const Tile* tiles;
  const R* Rin;
  R* Rout;
  const BCRec* bc;
  const int* fmap;
  R beta, den;         // den = 1 + 2 beta
  R inv;               // 1/den when den is a power of two (x/den == x*inv exactly), else 0
  // fused update (last direction)
  const R* T;
  const R* TO;
  R* Tout;
  const int* mat;
  Tab<R> tab;
  R rkc;
  int jam;
  int last;
  int irk;             // stage number of the step (state-check location)
  double* partials;
  int* status;
  const int* halt;
  // k_irs_xyz on the last stage: the next step's local dt from the new state
  // (Compute_dt of the next Explicit_Step, as k_dt computes it)
  int dtnext;
  double vnn;
  const double* invL2;
  const double* rect;  // non-null: sum 1/dl^2 from the per-axis factors instead of invL2
  R* dt;
  double* pmin;
};

// x / (1 + 2 beta): a multiplication by the exact reciprocal when the
// divisor is a power of two (beta = 0.5: den = 2), which gives the same
// IEEE result as the division; else the division.
template <typename R>
__device__ __forceinline__ R irs_div(R x, R den, R inv) {
  return inv != (R)0 ? x * inv : x / den;
}

// Residual_Smoothing_Blk along DIR (Lib_IRS.f90:43-134): two Jacobi sweeps
// r*_new(s) = (r(s) + beta (r*(s-1) + r*(s+1))) / (1 + 2 beta), r* = r
// initially, ghosts from Ghost_Residual_Extrapolation (:137-187).  The
// second sweep needs r(s-2..s+2) only, so each cell is computed directly.
// ------------------------------------------------------------------------
// Line-tridiagonal IRS (FUSS_IRS_TRIDIAG, no reference counterpart): per
// direction the exact solution of (1+2b) x_s - b (x_{s-1} + x_{s+1}) = r_s,
// s = 1..n, with the IRS ghosts x_0, x_{n+1} of Residual_Smoothing held
// fixed -- the fixed point of the reference's Jacobi sweeps
// (Lib_IRS.f90:63-89).  Thomas algorithm, one line per work-item; the
// elimination factors den_s = (1+2b) + b c_{s-1}, c_s = -b / den_s depend
// only on s and are tabulated per block and direction on the host (tri).
// Forward: d_1 = (r_1 + b x_0) / den_1, d_s = (r_s + b d_{s-1}) / den_s
// (r_n + b x_{n+1} on the last); back: x_n = d_n, x_s = d_s - c_s x_{s+1}.
// The d_s go to Rout and are overwritten by x_s on the way back.
// ------------------------------------------------------------------------
template <typename R>
struct TriArgs {
  const BlockInfo* blk;
  const Tile* tiles;
  const R* Rin;
  R* Rout;
  const BCRec* bc;
  const int* fmap;
  const double2* tri;  // (den_s, c_s) rows
  const int* trioff;   // per block and direction: first row of its table
  R beta;
  const int* halt;
};

// Ghost of a line end for Residual_Smoothing: the donor's R on a
// connection, else the line's own end value (Ghost_Residual_Extrapolation,
// Lib_IRS.f90:137-187).

// Two Jacobi sweeps at line position s of n (Lib_IRS.f90:63-89) from
// r(s-2..s+2); positions 0 and n+1 are the ghosts glo/ghi, which the sweeps
// leave unchanged (Residual_Smoothing_Blk smooths interior cells only).
template <typename R>
__device__ __forceinline__ R jacobi2(R rm2, R rm1, R r0, R rp1, R rp2, int s, int n, R glo, R ghi, R beta, R den,
                                     R inv) {
#pragma clang fp contract(off)
  const R am1 = s - 1 >= 1 ? rm1 : glo;
  const R ap1 = s + 1 <= n ? rp1 : ghi;
  const R am2 = s - 2 >= 1 ? rm2 : glo;
  const R ap2 = s + 2 <= n ? rp2 : ghi;
  const R sm = (s - 1 < 1) ? glo : irs_div(am1 + beta * (am2 + r0), den, inv);
  const R sp = (s + 1 > n) ? ghi : irs_div(ap1 + beta * (r0 + ap2), den, inv);
  return irs_div(r0 + beta * (sm + sp), den, inv);
}

// The same two sweeps with the division chosen at compile time (P2: den is
// a power of two and x * inv is the IEEE quotient), and jacobi2_in for a
// cell whose neighbours at distance 1 and 2 lie inside the line
// (3 <= s <= n-2), where jacobi2 selects no ghost: the same operations.
template <bool P2, typename R>
__device__ __forceinline__ R irs_q(R x, R den, R inv) {
  return P2 ? x * inv : x / den;
}

template <bool P2, typename R>
__device__ __forceinline__ R jacobi2_p(R rm2, R rm1, R r0, R rp1, R rp2, int s, int n, R glo, R ghi, R beta, R den,
                                       R inv) {
#pragma clang fp contract(off)
  const R am1 = s - 1 >= 1 ? rm1 : glo;
  const R ap1 = s + 1 <= n ? rp1 : ghi;
  const R am2 = s - 2 >= 1 ? rm2 : glo;
  const R ap2 = s + 2 <= n ? rp2 : ghi;
  const R sm = (s - 1 < 1) ? glo : irs_q<P2>(am1 + beta * (am2 + r0), den, inv);
  const R sp = (s + 1 > n) ? ghi : irs_q<P2>(ap1 + beta * (r0 + ap2), den, inv);
  return irs_q<P2>(r0 + beta * (sm + sp), den, inv);
}

template <bool P2, typename R>
__device__ __forceinline__ R jacobi2_in(R rm2, R rm1, R r0, R rp1, R rp2, R beta, R den, R inv) {
#pragma clang fp contract(off)
  const R sm = irs_q<P2>(rm1 + beta * (rm2 + r0), den, inv);
  const R sp = irs_q<P2>(rp1 + beta * (r0 + rp2), den, inv);
  return irs_q<P2>(r0 + beta * (sm + sp), den, inv);
}

constexpr int XY_TJ = 16;                 // rows of an x-y IRS tile
constexpr int XY_RJ = XY_TJ + 4, XY_RI = TILE_X + 4;

// Row of T in a 1 K table (co_Kappa etc., Lib_Solid.f90:19-23: Tm = int(T)).
// Outside [Ti1, Ti2) *bad is set (the step then fails with FUSS_ETABLE) and
// T is clamped to [Ti1, Ti2 - 1] before the conversion, so the row stays
// inside the table and the conversion never sees a NaN or an out-of-range
// value (fmax of a NaN returns the bound).  With Ti1 >= 0, int(T) >= Ti2
// exactly when T >= Ti2.
// tab_row's row as a floating value (the integer part of the clamped T):
// T - row is the interpolation's fraction without a conversion back from the
// integer, and (int) of it addresses the row
template <typename R>
__device__ __forceinline__ R tab_rowf(R T, int Ti1, int Ti2, int* bad) {
  if (!(T >= (R)Ti1 && T < (R)Ti2)) *bad = 1;
  return trunc(fmin(fmax(T, (R)Ti1), (R)(Ti2 - 1)));
}
template <typename R>
__device__ __forceinline__ int tab_row(R T, int Ti1, int Ti2, int* bad) {
  if (!(T >= (R)Ti1 && T < (R)Ti2)) *bad = 1;
  return (int)fmin(fmax(T, (R)Ti1), (R)(Ti2 - 1));
}

// Table row of T clamped to [Ti1, Ti2 - 1]: the hardware conversion
// saturates (and maps NaN to 0), so with Ti1 >= 0 this is
// trunc(fmin(fmax(T, Ti1), Ti2 - 1)) of tab_rowf as an integer.
// (hi is held in a VGPR: one instruction may read one scalar register)
__device__ __forceinline__ int med3_i32(int c, int lo, int hi) {
  int r;
  asm("v_med3_i32 %0, %1, %2, %3" : "=v"(r) : "v"(c), "s"(lo), "v"(hi));
  return r;
}
// median of (c, 0, hi) for a uniform hi (no VGPR for the bound)
__device__ __forceinline__ int med3_i32_0s(int c, int hi) {
  int r;
  asm("v_med3_i32 %0, %1, 0, %2" : "=v"(r) : "v"(c), "s"(hi));
  return r;
}
__device__ __forceinline__ int row_sat(double T, int Ti1, int Ti2m1) {
  int c;
  asm("v_cvt_i32_f64 %0, %1" : "=v"(c) : "v"(T));
  return med3_i32(c, Ti1, Ti2m1);
}
__device__ __forceinline__ int row_sat(float T, int Ti1, int Ti2m1) {
  int c;
  asm("v_cvt_i32_f32 %0, %1" : "=v"(c) : "v"(T));
  return med3_i32(c, Ti1, Ti2m1);
}
// a uniform value copied to a VGPR once
__device__ __forceinline__ int to_vgpr(int s) {
  int v;
  asm volatile("v_mov_b32 %0, %1" : "=v"(v) : "s"(s));
  return v;
}

// An opaque statement: a branch on a uniform condition that contains it
// stays a scalar branch (it is not speculated into a select).
__device__ __forceinline__ void keep_branch() { asm volatile(""); }

// x - floor(x): for x >= 0 the interpolation fraction x - trunc(x) of a
// table row (exact: the integer part's subtraction loses nothing)
__device__ __forceinline__ double frac_of(double x) { return __builtin_amdgcn_fract(x); }
__device__ __forceinline__ float frac_of(float x) { return __builtin_amdgcn_fractf(x); }
