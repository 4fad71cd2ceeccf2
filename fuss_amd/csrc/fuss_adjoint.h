// fuss_adjoint.h -- discrete adjoint of the time-accurate explicit step.
//
// Forward map of one step (Explicit_Step with a global dt, primitive
// variables, no IRS; src/lib/numerics/time/explicit/Mod_Explicit.f90:10-111):
//   dt     = min(1e5, min_c 0.5 vnn rho cp / (kappa sum 1/dl^2))     (Mod_dt.f90:26-84)
//   T_1    = T^n
//   T_s+1  = T^n + c_s (T_s - T^n + res_s)        (SSP stages, Lib_RK.f90:15-26)
//          = T^n + a_s res_s                      (four-stage extension)
//   res_s  = (-R(T_s)/V + qvol) dt / cpe(T_s),  cpe = T cp rho' + rho cp + T rho cp'
//   R      : interior faces R_l -= F, R_r += F with F = kappa(Tbar) W (T_r - T_l),
//            kappa of the left cell; BC faces R -= v_e   (Mod_Fluxes.f90, Mod_BC_Fluxes.f90)
//   T^n+1  = T_S+1
// Parameters: the per-entry wall values of 301 entries (v = modfm2 A qw) and a
// scale c of the conductivity table (kappa -> c kappa, which makes every
// flux, the 302/303 wall fluxes and dt depend on c).
//
// The reverse sweep visits the stages backwards.  k_adj_stage computes, in
// gather form (each cell reads the adjoints of its neighbours, no atomics),
// the adjoint of the stage input T_s from the adjoint of its output, adds
// the stage's share to the adjoint of T^n, and emits per-work-group partials
// of dL/d(dt) and dL/dc and per-entry dL/dqw.  The dt path (the global min)
// is differentiated through its first minimising cell.  Scope: orthogonal
// blocks (no tangential weights), wall types 0/301/302/303, no connections.
#pragma once

namespace fuss {

template <typename S>
struct AdjArgs {
  const BlockInfo* blk;
  const Tile* tiles;
  const double* T;      // stage input T_s (ghosts filled)
  const double* qvol;
  const double* vol;
  const int* mat;
  const double* W[3];   // normal face weights W[3D+D]
  const BCRec* bc;
  const int* fmap;
  const double* bcv;    // the stage's BC values v_e
  Tab<double> tab;
  const double* dtg;    // dt of the step (device)
  double coef;          // c_s (SSP) or a_s (four-stage)
  int jam;
  double kscale;
  const S* lout;        // adjoint of T_{s+1}
  S* lin;               // adjoint of T_s (written)
  S* lTO;               // adjoint of T^n (accumulated)
  double* part;         // 2 per work-group: dL/d(dt), dL/dc
  double* gqw;          // per BC entry, accumulated dL/dqw
};

// rho cp T-derivative form of Compute_Prim_Residual's denominator and its slope
// (within one 1 K row rho' and cp' are constant).
__device__ __forceinline__ void adj_cpe(const Tab<double>& tab, int m, double T, double* cpe, double* dcpe) {
  double Tc = T;
  if (!(Tc >= (double)tab.Ti1 && Tc < (double)tab.Ti2)) Tc = Tc < (double)tab.Ti1 || Tc != Tc ? (double)tab.Ti1 : (double)(tab.Ti2 - 1);
  const int Tm = (int)Tc;
  const double4 v = tab.pr4 ? reinterpret_cast<const double4*>(tab.pr4)[(long)m * tab.nT + (Tm - tab.Ti1)]
                            : double4{0, 0, 0, 0};
  const double rho = v.x + (Tc - (double)Tm) * v.y;
  const double cp = v.z + (Tc - (double)Tm) * v.w;
  const double drho = v.y, dcp = v.w;
  *cpe = T * cp * drho + rho * cp + T * rho * dcp;
  *dcpe = 2.0 * (cp * drho + rho * dcp) + 2.0 * T * drho * dcp;
}

__device__ __forceinline__ void adj_kap(const Tab<double>& tab, int m, double T, double* k, double* dk) {
  double Tc = T;
  if (!(Tc >= (double)tab.Ti1 && Tc < (double)tab.Ti2)) Tc = Tc < (double)tab.Ti1 || Tc != Tc ? (double)tab.Ti1 : (double)(tab.Ti2 - 1);
  const int Tm = (int)Tc;
  const double2 v = reinterpret_cast<const double2*>(tab.kap2)[(long)m * tab.nT + (Tm - tab.Ti1)];
  *k = v.x + (Tc - (double)Tm) * v.y;
  *dk = v.y;
}

template <int NDIR, typename S>
__global__ __launch_bounds__(TILE_T) void k_adj_stage(AdjArgs<S> a) {
  const int wg = blockIdx.x;
  const Tile t = a.tiles[wg];
  const BlockInfo B = a.blk[t.b];
  const int i = t.i0 + (threadIdx.x % TILE_X), j = t.j0 + (threadIdx.x / TILE_X), k = t.k0;
  double pdt = 0.0, pc = 0.0;
  if (i <= B.ni && j <= B.nj) {
    const long p = B.base + (i - 1) + (long)(j - 1) * B.sy + (long)(k - 1) * B.sz;
    const double dt = *a.dtg;
    const double* __restrict__ T = a.T;
    // adjoint of R at a cell x of the block
    auto lamR = [&](long x) -> double {
      double cpe, dcpe;
      adj_cpe(a.tab, a.mat[x], T[x], &cpe, &dcpe);
      return -(a.coef * (double)a.lout[x]) * (dt / cpe) / a.vol[x];
    };
    const double lRc = lamR(p);
    double acc = 0.0, Rv = 0.0;
    const int ni[3] = {B.ni, B.nj, B.nk};
    const int ci[3] = {i, j, k};
    for (int D = 0; D < NDIR; D++) {
      const long s = D == 0 ? 1 : (D == 1 ? (long)B.sy : B.sz);
      if (ci[D] > 1) {  // face on the min side: this cell is the right one
        const long q = p - s;
        double kf, dk;
        adj_kap(a.tab, a.mat[q], 0.5 * (T[q] + T[p]), &kf, &dk);
        const double w = a.W[D][q], G = T[p] - T[q];
        const double F = kf * w * G;
        Rv += F;
        acc += (lRc - lamR(q)) * (0.5 * dk * w * G + kf * w);
      }
      if (ci[D] < ni[D]) {  // face on the max side: this cell is the left one
        const long q = p + s;
        double kf, dk;
        adj_kap(a.tab, a.mat[p], 0.5 * (T[p] + T[q]), &kf, &dk);
        const double w = a.W[D][p], G = T[q] - T[p];
        const double F = kf * w * G;
        Rv -= F;
        const double lF = lamR(q) - lRc;
        acc += lF * (0.5 * dk * w * G - kf * w);
        pc += lF * F / a.kscale;
      }
    }
    // BC faces of the cell, R -= v_e
    const int fa = B.face_active;
    if (fa) {
      const double Tc = T[p];
      for (int f = 1; f <= 2 * NDIR; f++) {
        if (!(fa & (1 << (f - 1)))) continue;
        const int d = (f - 1) / 2;
        const bool lo = (f & 1) != 0;
        if (ci[d] != (lo ? 1 : ni[d])) continue;
        const int e = a.fmap[face_map_index<double>(B, f, i, j, k)];
        if (e < 0) continue;
        const BCRec& r = a.bc[e];
        Rv -= a.bcv[e];
        const double lv = -lRc;
        const int modfm2 = 1 - 2 * (r.f % 2);
        const double* M = r.M;
        const double mn = M[r.dir - 1] * r.N[0] + M[r.dir - 1 + 3] * r.N[1] + M[r.dir - 1 + 6] * r.N[2];
        if (r.type == 301) {
          if (r.series < 0) a.gqw[e] += lv * (double)modfm2 * r.A;
        } else if (r.type == 302) {
          double kc, dkc;
          adj_kap(a.tab, a.mat[p], Tc, &kc, &dkc);
          const double v = a.bcv[e];
          acc += lv * (-2.0 * r.A * mn * (dkc * (Tc - r.Tw) + kc));
          pc += lv * v / a.kscale;
        } else if (r.type == 303) {
          double kc, dkc;
          adj_kap(a.tab, a.mat[p], Tc, &kc, &dkc);
          const double H = 1.0 / (2.0 * kc * mn) + 1.0 / r.hconv;
          const double g = r.A * (r.Tref - Tc) / (H * H);
          acc += lv * (-r.A / H + g * dkc / (2.0 * mn * kc * kc));
          pc += lv * g / (2.0 * mn * kc * a.kscale);
        }
      }
    }
    // residual path: res = (-R/V + qvol) dt / cpe(T_s)
    double cpe, dcpe;
    adj_cpe(a.tab, a.mat[p], T[p], &cpe, &dcpe);
    const double Q = -Rv / a.vol[p] + a.qvol[p];
    const double lres = a.coef * (double)a.lout[p];
    acc += lres * (-Q * dt * dcpe / (cpe * cpe));
    pdt = lres * Q / cpe;
    // direct terms of the RK combination
    const double lo_ = (double)a.lout[p];
    if (a.jam) {
      a.lTO[p] = (S)((double)a.lTO[p] + lo_);
    } else {
      acc += a.coef * lo_;
      a.lTO[p] = (S)((double)a.lTO[p] + (1.0 - a.coef) * lo_);
    }
    a.lin[p] = (S)acc;
  }
  // per-work-group sums of the two scalar adjoints
  __shared__ double sm[2][TILE_T / 64];
  for (int o = 32; o > 0; o >>= 1) {
    pdt += __shfl_xor(pdt, o, 64);
    pc += __shfl_xor(pc, o, 64);
  }
  if ((threadIdx.x & 63) == 0) {
    sm[0][threadIdx.x >> 6] = pdt;
    sm[1][threadIdx.x >> 6] = pc;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    double s0 = sm[0][0], s1 = sm[1][0];
    for (int q = 1; q < TILE_T / 64; q++) {
      s0 += sm[0][q];
      s1 += sm[1][q];
    }
    a.part[2 * wg] = s0;
    a.part[2 * wg + 1] = s1;
  }
}

// lam = lam + lTO on interior cells; lTO = 0.
template <typename S>
__global__ __launch_bounds__(TILE_T) void k_adj_combine(const BlockInfo* __restrict__ blk, const Tile* __restrict__ tiles,
                                                        S* __restrict__ lam, S* __restrict__ lTO) {
  const Tile t = tiles[blockIdx.x];
  const BlockInfo B = blk[t.b];
  const int i = t.i0 + (threadIdx.x % TILE_X), j = t.j0 + (threadIdx.x / TILE_X), k = t.k0;
  if (i > B.ni || j > B.nj) return;
  const long p = B.base + (i - 1) + (long)(j - 1) * B.sy + (long)(k - 1) * B.sz;
  lam[p] = (S)((double)lam[p] + (double)lTO[p]);
  lTO[p] = (S)0;
}

// Deterministic sums of the (dt, c) partials: pass 1 per RED_CHUNK pairs.
__global__ __launch_bounds__(256) void k_adj_sum2(const double* __restrict__ part, int n, double* __restrict__ out) {
  __shared__ double sm[2][256];
  const int lo = blockIdx.x * RED_CHUNK, hi = min(n, lo + RED_CHUNK);
  double a = 0.0, b = 0.0;
  for (int q = lo + threadIdx.x; q < hi; q += 256) {
    a += part[2 * q];
    b += part[2 * q + 1];
  }
  sm[0][threadIdx.x] = a;
  sm[1][threadIdx.x] = b;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if (threadIdx.x < w) {
      sm[0][threadIdx.x] += sm[0][threadIdx.x + w];
      sm[1][threadIdx.x] += sm[1][threadIdx.x + w];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    out[2 * blockIdx.x] = sm[0][0];
    out[2 * blockIdx.x + 1] = sm[1][0];
  }
}
// Pass 2: acc[0] += sum dt-adjoints, acc[1] += sum c-adjoints.
__global__ void k_adj_acc2(const double* __restrict__ in, int n, double* __restrict__ acc) {
  if (threadIdx.x != 0) return;
  double a = 0.0, b = 0.0;
  for (int q = 0; q < n; q++) {
    a += in[2 * q];
    b += in[2 * q + 1];
  }
  acc[0] += a;
  acc[1] += b;
}

// First cell (arena order) whose local dt equals the step's dt: per tile.
__global__ __launch_bounds__(TILE_T) void k_adj_argmin(const BlockInfo* __restrict__ blk, const Tile* __restrict__ tiles,
                                                       const double* __restrict__ dtl, const double* __restrict__ dtg,
                                                       int* __restrict__ out) {
  const Tile t = tiles[blockIdx.x];
  const BlockInfo B = blk[t.b];
  const int i = t.i0 + (threadIdx.x % TILE_X), j = t.j0 + (threadIdx.x / TILE_X), k = t.k0;
  int best = 0x7fffffff;
  if (i <= B.ni && j <= B.nj) {
    const long p = B.base + (i - 1) + (long)(j - 1) * B.sy + (long)(k - 1) * B.sz;
    if (dtl[p] == *dtg) best = (int)p;
  }
  for (int o = 32; o > 0; o >>= 1) best = min(best, __shfl_xor(best, o, 64));
  __shared__ int sm[TILE_T / 64];
  if ((threadIdx.x & 63) == 0) sm[threadIdx.x >> 6] = best;
  __syncthreads();
  if (threadIdx.x == 0) {
    int v = sm[0];
    for (int q = 1; q < TILE_T / 64; q++) v = min(v, sm[q]);
    out[blockIdx.x] = v;
  }
}

// dt path of one step: dL/d(dt) (acc[0]) onto T^n at the minimising cell and
// onto c (dt ~ 1/c); acc[0] is then cleared for the next step.
__global__ void k_adj_dtpath(const int* __restrict__ amin, int ntile, const double* __restrict__ Tn,
                             const int* __restrict__ mat, const double* __restrict__ dtg, Tab<double> tab,
                             double kscale, double* __restrict__ lam_d, float* __restrict__ lam_f,
                             double* __restrict__ acc) {
  if (threadIdx.x != 0) return;
  const double dt = *dtg;
  const double ldt = acc[0];
  acc[0] = 0.0;
  if (!(dt < 1e5)) return;  // capped: dt does not depend on the state
  int p = 0x7fffffff;
  for (int q = 0; q < ntile; q++) p = min(p, amin[q]);
  if (p == 0x7fffffff) return;
  const double T = Tn[p];
  const int m = mat[p];
  double kc, dk;
  adj_kap(tab, m, T, &kc, &dk);
  double Tc = T;
  if (!(Tc >= (double)tab.Ti1 && Tc < (double)tab.Ti2)) Tc = Tc < (double)tab.Ti1 || Tc != Tc ? (double)tab.Ti1 : (double)(tab.Ti2 - 1);
  const int Tm = (int)Tc;
  const double4 v = reinterpret_cast<const double4*>(tab.pr4)[(long)m * tab.nT + (Tm - tab.Ti1)];
  const double rho = v.x + (Tc - (double)Tm) * v.y, cp = v.z + (Tc - (double)Tm) * v.w;
  const double ddt = dt * (v.y / rho + v.w / cp - dk / kc);
  if (lam_d) lam_d[p] += ldt * ddt;
  else lam_f[p] = (float)((double)lam_f[p] + ldt * ddt);
  acc[1] += -ldt * dt / kscale;
}

// lam[probe q] += g[q] (one work-item: probes may share a cell).
template <typename S>
__global__ void k_adj_seed(S* __restrict__ lam, const long* __restrict__ idx, const double* __restrict__ g, int n) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  for (int q = 0; q < n; q++) lam[idx[q]] = (S)((double)lam[idx[q]] + g[q]);
}

__global__ void k_probe_rec(const double* __restrict__ T, const long* __restrict__ idx, double* __restrict__ out,
                            int n) {
  const int q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q < n) out[q] = T[idx[q]];
}

// Interior of lam (arena) to a compact i-fastest array per block.
template <typename S>
__global__ __launch_bounds__(TILE_T) void k_adj_extract(const BlockInfo* __restrict__ blk, const Tile* __restrict__ tiles,
                                                        const long* __restrict__ coff, const S* __restrict__ lam,
                                                        double* __restrict__ out) {
  const Tile t = tiles[blockIdx.x];
  const BlockInfo B = blk[t.b];
  const int i = t.i0 + (threadIdx.x % TILE_X), j = t.j0 + (threadIdx.x / TILE_X), k = t.k0;
  if (i > B.ni || j > B.nj) return;
  const long p = B.base + (i - 1) + (long)(j - 1) * B.sy + (long)(k - 1) * B.sz;
  out[coff[t.b] + (i - 1) + (long)B.ni * ((j - 1) + (long)B.nj * (k - 1))] = (double)lam[p];
}

// Level time and iteration before a step (forward of the adjoint).
__global__ void k_clock_rec(const double* __restrict__ clk, const int* __restrict__ cnt, double* t, int* it) {
  if (threadIdx.x == 0 && blockIdx.x == 0) {
    *t = clk[0];
    *it = cnt[0];
  }
}

// Wall values of selected BC entries (fuss_set_bc_values): column 0 qw,
// 1 Tw, 2 hconv, 4 Tref, 5 eps.
__global__ void k_set_bcval(BCRec* __restrict__ bc, const int* __restrict__ idx, const double* __restrict__ v, int n,
                            int col) {
  const int q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= n) return;
  BCRec& r = bc[idx[q]];
  if (col == 0) r.qw = v[q];
  else if (col == 1) r.Tw = v[q];
  else if (col == 2) r.hconv = v[q];
  else if (col == 4) r.Tref = v[q];
  else if (col == 5) r.eps = v[q];
}

template <typename S>
__global__ void k_adj_fill(S* __restrict__ p, long n) {
  const long g = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (g < n) p[g] = (S)0;
}

}  // namespace fuss
