// persist.hip -- persistent explicit steps for latency-bound levels.
//
// One cooperative launch runs up to `maxcount` Explicit_Steps of a level
// (Mod_Explicit.f90:10-111) with the stop tests of the step loop
// (:81-90) on the device: the work-groups march through the phases of a
// step -- local dt of their tiles, then per RK stage the fused interior +
// BC flux + newstate pass (with the BC values evaluated in place and the
// layer-1 mirror ghosts of the new state written with it) -- separated by
// grid-wide barriers, then every work-group reduces the residual partials
// itself (the step loop's fixed order) and applies the same stop test.
//
// The arithmetic of every phase is the step loop's (same device functions,
// same tile partition of the partial sums), so the states, residuals and
// clocks equal fuss_run's step by step.  Scope: levels without connections
// (101/102) and without IRS; the caller checks.
//
// Grid barriers: an arrival counter and a release counter in device memory,
// agent-scope fences around them (the XCDs' L2 caches are written back and
// invalidated there).  The last work-group to arrive snapshots the status
// word, so every work-group takes the same stop decision.  Every wait is
// bounded: past PK_SPIN_LIMIT polls the work-group raises the abort word,
// which ends every other wait, and the launch reports the failure.
#include "fuss_host.h"

namespace fuss {

constexpr unsigned PK_SPIN_LIMIT = 1u << 20;


// Grid-wide barrier; returns false when a wait timed out (or another
// work-group's did).  *snap receives the status word as the last arriving
// work-group read it.
template <typename R>
__device__ bool grid_sync(const PArgs<R>& a, unsigned& gen, int* snap) {
  __shared__ int sh[2];
  // every wave's stores have reached the L2 of its XCD (waves of a
  // work-group share one CU), so thread 0's agent-scope release below
  // writes them back with its own
  __builtin_amdgcn_s_waitcnt(0);
  __syncthreads();
  if (threadIdx.x == 0) {
    const unsigned target = (gen + 1) * gridDim.x;
    int ok = 1;
    __threadfence();
    const unsigned old = atomicAdd(&a.bar[0], 1u);
    if (old == target - 1) {
      __threadfence();
      __hip_atomic_store(&a.ctl[1], __hip_atomic_load(a.sa.status, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT),
                         __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      __hip_atomic_store(&a.bar[1], gen + 1, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
    } else {
      unsigned spins = 0;
      while (__hip_atomic_load(&a.bar[1], __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT) < gen + 1) {
        if (__hip_atomic_load(&a.ctl[0], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) || ++spins > PK_SPIN_LIMIT) {
          __hip_atomic_store(&a.ctl[0], 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
          ok = 0;
          break;
        }
        __builtin_amdgcn_s_sleep(1);
      }
    }
    __threadfence();  // the other work-groups' results, after the release
    sh[0] = ok;
    sh[1] = __hip_atomic_load(&a.ctl[1], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  }
  __syncthreads();  // thread 0's acquire invalidated the CU's and XCD's caches for every wave
  gen++;
  *snap = sh[1];
  return sh[0] != 0;
}

// dtglobal = min(1e5, min over the tiles' minima): k_dt_final's partition
// and tree, in every work-group (256 threads).
__device__ double pk_min(const double* __restrict__ in, int n) {
  __shared__ double sm[256];
  double v = 1e300;
  for (int q = threadIdx.x; q < n; q += 256) v = fmin(v, __hip_atomic_load(in + q, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT));
  sm[threadIdx.x] = v;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if (threadIdx.x < w) sm[threadIdx.x] = fmin(sm[threadIdx.x], sm[threadIdx.x + w]);
    __syncthreads();
  }
  const double r = fmin(1e5, sm[0]);
  __syncthreads();
  return r;
}

// Sum (T - TO)^2 over the per-tile partials: k_reduce_sum1's chunks, thread
// partition and trees, in every work-group (256 threads).
__device__ double pk_sum(const double* __restrict__ in, int n) {
  __shared__ double sm[256];
  __shared__ double cs[256];
  cs[threadIdx.x] = 0.0;
  const int nb = (n + 2047) / 2048;
  for (int b = 0; b < nb; b++) {
    const int lo = b * 2048, hi = min(n, lo + 2048);
    double v = 0.0;
    for (int q = lo + threadIdx.x; q < hi; q += 256)
      v = v + __hip_atomic_load(in + q, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    sm[threadIdx.x] = v;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
      if (threadIdx.x < w) sm[threadIdx.x] = sm[threadIdx.x] + sm[threadIdx.x + w];
      __syncthreads();
    }
    if (threadIdx.x == 0) cs[b] = 0.0 + sm[0];
    __syncthreads();
  }
  for (int w = 128; w > 0; w >>= 1) {
    if (threadIdx.x < w) cs[threadIdx.x] = cs[threadIdx.x] + cs[threadIdx.x + w];
    __syncthreads();
  }
  const double r = cs[0];
  __syncthreads();
  return r;
}

// The BC value of face f of cell (i, j, k) (0 without an entry), as the
// surface pass stores it for apply_bc_faces.
template <int NDIR, typename R>
__device__ __forceinline__ R pk_face_value(const StageArgs<R>& a, const BCHot* hot, const BlockInfo& B, int f, int i,
                                           int j, int k, double time) {
  const int e = a.fmap[face_map_index<R>(B, f, i, j, k)];
  return e >= 0 ? bc_entry_value<NDIR>(a, hot, e, time) : (R)0;
}

template <typename R>
__device__ __forceinline__ void pk_ghost(const StageArgs<R>& a, const int2* gmap, const BlockInfo& B, int f, int i,
                                         int j, int k, R v) {
  if (B.fmap_off[f - 1] < 0) return;  // no entries on this face
  const int e = a.fmap[face_map_index<R>(B, f, i, j, k)];
  if (e >= 0) {
    const int2 g = gmap[e];
    if (g.x >= 0) stu(a.Tout, (unsigned)g.x, v);
  }
}

// One tile of k_stage (update mode) with the BC values evaluated here and
// the mirror ghosts of the new state written (same operations and order).
template <int NDIR, bool TANG, bool PRIM, typename R>
__device__ __forceinline__ void pk_stage_tile(const StageArgs<R>& a, const BCHot* hot, const int2* gmap, int ghosts,
                                              int wt, double time, R dtg) {
  const Tile t = a.tiles[wt];
  const BlockInfo B = a.blk[t.b];
  const int i = t.i0 + (threadIdx.x % TILE_X), j = t.j0 + (threadIdx.x / TILE_X), k = t.k0;
  double rsq = 0.0;
  if (i <= B.ni && j <= B.nj) {
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
    {
#pragma clang fp contract(off)
      // BC_Fluxes contributions in face order (apply_bc_faces)
      const int fa = B.face_active;
      if (fa) {
        if ((fa & 1) && i == 1) Rv = Rv - pk_face_value<NDIR>(a, hot, B, 1, i, j, k, time);
        if ((fa & 2) && i == B.ni) Rv = Rv - pk_face_value<NDIR>(a, hot, B, 2, i, j, k, time);
        if ((fa & 4) && j == 1) Rv = Rv - pk_face_value<NDIR>(a, hot, B, 3, i, j, k, time);
        if ((fa & 8) && j == B.nj) Rv = Rv - pk_face_value<NDIR>(a, hot, B, 4, i, j, k, time);
        if ((fa & 16) && k == 1) Rv = Rv - pk_face_value<NDIR>(a, hot, B, 5, i, j, k, time);
        if ((fa & 32) && k == B.nk) Rv = Rv - pk_face_value<NDIR>(a, hot, B, 6, i, j, k, time);
      }
    }
    const R T = a.T[p];
    const int m = a.mat[p];
    const R dt = a.use_dtg ? dtg : a.dt[p];
    int bad = 0;
    const R res = scaled_residual<PRIM>(a.tab, m, T, Rv, a.vol[p], a.qvol[p], dt, &bad);
    const R TO = a.TO[p];
    const R Tn = new_state<PRIM>(a.tab, m, T, TO, res, a.rkc, a.jam, &bad);
    a.Tout[p] = Tn;
    if (Tn != Tn || Tn < (R)0) flag_state(a.status, a.irk, t.b, i, j, k);
    const double d = (double)Tn - (double)TO;
    rsq = d * d;
    if (bad) flag_error(a.status, FUSS_ETABLE, t.b, i, j, k);
    if (ghosts) {
      // Fill_Ghost_Cell's mirror (layer 1) of the new state for the next stage
      if (i == 1) pk_ghost(a, gmap, B, 1, i, j, k, Tn);
      if (i == B.ni) pk_ghost(a, gmap, B, 2, i, j, k, Tn);
      if (j == 1) pk_ghost(a, gmap, B, 3, i, j, k, Tn);
      if (j == B.nj) pk_ghost(a, gmap, B, 4, i, j, k, Tn);
      if (k == 1) pk_ghost(a, gmap, B, 5, i, j, k, Tn);
      if (k == B.nk) pk_ghost(a, gmap, B, 6, i, j, k, Tn);
    }
  }
  if (a.last) block_sum(rsq, a.partials, wt);
  __syncthreads();  // block_sum's LDS is reused by the next tile
}

// ------------------------------------------------------------------------
// Fused steps of latency-bound 2-D levels without IRS and connections (the
// step loop's path, one launch per stage): each stage kernel evaluates its
// cells' BC values in place (no surface pass) and writes the mirror ghosts
// of the new state (no ghost pass before the next stage); on steady steps
// the first stage also computes the local dt (no dt pass), and one kernel
// at the end of the step reduces the residual partials and the dt minima
// and advances the clock.  Same arithmetic and order as the split passes.
// ------------------------------------------------------------------------
template <int NDIR, bool TANG, bool PRIM, bool DT, typename R>
__global__ __launch_bounds__(TILE_T) void k_stage_fused(StageArgs<R> a, DtArgs<R> da, const BCHot* __restrict__ hot,
                                                        const int2* __restrict__ gmap, int ghosts) {
  if (a.halt && *a.halt) return;
  const int wg = xcd_remap(blockIdx.x, gridDim.x);
  if (DT) dt_tile(da, wg, da.cnt[0] + 1);  // Compute_dt at T^n, this stage's input
  pk_stage_tile<NDIR, TANG, PRIM>(a, hot, gmap, ghosts, wg, *a.clk, a.use_dtg ? (R)(*a.dtg) : (R)0);
}

// Sum (T-TO)^2 (k_reduce_sum1) and, when the first stage computed dt,
// k_dt_final's minimum, dtglobal and clock advance, in one work-group.
__global__ __launch_bounds__(256) void k_fused_end(const double* __restrict__ part, int n, const double* __restrict__ pmin,
                                                   int dt_here, double* scal, double* clk, int* cnt, int time_accurate) {
  if (cnt[1]) return;
  const double rs = pk_sum(part, n);
  double m = 0.0;
  if (dt_here) m = pk_min(pmin, n);
  if (threadIdx.x == 0) {
    scal[2] = rs;
    if (dt_here) {
      const double t = clk[0];
      const double tn = time_accurate ? t + m : t;
      scal[0] = m;
      scal[1] = tn;
      clk[0] = tn;
      cnt[0] += 1;
    }
  }
}

template <typename R>
void launch_stage_fused(fuss_ctx* c, Level& L, const StageArgs<R>& a, const DtArgs<R>& da, int ndir, bool tang,
                        bool prim, bool dt_here, bool ghosts) {
  const unsigned g = (unsigned)L.htiles.size();
  const int gh = ghosts ? 1 : 0;
#define FUSS_SF(N, TG, P, D) k_stage_fused<N, TG, P, D, R><<<g, TILE_T, 0, c->stream>>>(a, da, L.bch.p, L.gmap.p, gh)
  if (ndir == 3) {
    if (tang) { if (prim) { dt_here ? FUSS_SF(3, true, true, true) : FUSS_SF(3, true, true, false); }
                else { dt_here ? FUSS_SF(3, true, false, true) : FUSS_SF(3, true, false, false); } }
    else { if (prim) { dt_here ? FUSS_SF(3, false, true, true) : FUSS_SF(3, false, true, false); }
           else { dt_here ? FUSS_SF(3, false, false, true) : FUSS_SF(3, false, false, false); } }
  } else {
    if (tang) { if (prim) { dt_here ? FUSS_SF(2, true, true, true) : FUSS_SF(2, true, true, false); }
                else { dt_here ? FUSS_SF(2, true, false, true) : FUSS_SF(2, true, false, false); } }
    else { if (prim) { dt_here ? FUSS_SF(2, false, true, true) : FUSS_SF(2, false, true, false); }
           else { dt_here ? FUSS_SF(2, false, false, true) : FUSS_SF(2, false, false, false); } }
  }
#undef FUSS_SF
}

void launch_fused_end(fuss_ctx* c, Level& L, int n, bool dt_here, int time_accurate) {
  k_fused_end<<<1, 256, 0, c->stream>>>(L.partials.p, n, L.pmin.p, dt_here ? 1 : 0, c->scal.p, L.clk.p, L.cnt.p,
                                        time_accurate);
}

template void launch_stage_fused<double>(fuss_ctx*, Level&, const StageArgs<double>&, const DtArgs<double>&, int, bool,
                                         bool, bool, bool);
template void launch_stage_fused<float>(fuss_ctx*, Level&, const StageArgs<float>&, const DtArgs<float>&, int, bool,
                                        bool, bool, bool);

template <int NDIR, bool TANG, bool PRIM, typename R>
__global__ __launch_bounds__(TILE_T) void k_persist(PArgs<R> a) {
  const int G = gridDim.x, wg = blockIdx.x;
  unsigned gen = 0;
  double time = a.time0, dtglobal = 0.0, res_sq = 0.0;
  int cur = a.cur0, done = 0, snap = 0, stop = 0;
  for (int s = 0; s < a.maxcount; s++) {
    const int iter = a.iter0 + s + 1;
    R* const Tn = a.Tb[cur];
    // tile minima and residual partials alternate between two halves by
    // step parity: a work-group still reducing step s never sees step s+1's
    // values (a barrier separates step s from step s+2)
    double* const pmin = a.pmin2 + (s & 1) * a.ntile;
    double* const part = a.part2 + (s & 1) * a.ntile;
    DtArgs<R> da = a.da;
    da.T = Tn;
    da.pmin = pmin;
    for (int t = wg; t < a.ntile; t += G) dt_tile(da, t, iter);
    R dtg = 0;
    if (a.time_accurate) {
      // Set_Global_dt: every stage uses the global minimum at t^n + dt
      if (!grid_sync(a, gen, &snap)) return;
      dtglobal = pk_min(pmin, a.ntile);
      time = time + dtglobal;
      dtg = (R)dtglobal;
    }
    const int a1 = (cur + 1) % 3, a2 = (cur + 2) % 3;
    R* in = Tn;
    for (int irk = 1; irk <= a.nrk; irk++) {
      R* out = (irk % 2 == 1) ? a.Tb[a1] : a.Tb[a2];
      StageArgs<R> sa = a.sa;
      sa.T = in;
      sa.TO = Tn;
      sa.Tout = out;
      sa.rkc = (R)a.rkc[irk - 1];
      sa.jam = a.jam;
      sa.last = irk == a.nrk && a.compute_residual;
      sa.irk = irk;
      sa.partials = part;
      for (int t = wg; t < a.ntile; t += G)
        pk_stage_tile<NDIR, TANG, PRIM>(sa, a.hot, a.gmap, a.ghosts, t, time, dtg);
      if (!grid_sync(a, gen, &snap)) return;
      in = out;
    }
    if (!a.time_accurate) dtglobal = pk_min(pmin, a.ntile);
    res_sq = a.compute_residual ? pk_sum(part, a.ntile) : 0.0;
    cur = (a.nrk % 2 == 1) ? a1 : a2;
    done = s + 1;
    const double res = sqrt(res_sq);
    if (wg == 0 && threadIdx.x == 0) {
      if (a.hres) a.hres[s] = res;
      if (a.htime) a.htime[s] = time;
    }
    // k_step_end's stop tests, on values every work-group holds alike
    stop = snap || res <= a.thr || time >= a.tthr;
    if (stop || done >= a.maxcount) break;
  }
  if (wg == 0 && threadIdx.x == 0) {
    a.scal[0] = dtglobal;
    a.scal[1] = time;
    a.scal[2] = res_sq;
    a.clk[0] = time;
    a.cnt[0] = a.iter0 + done;
    a.ctl[2] = done;
    a.ctl[3] = cur;
    a.ctl[4] = stop;
  }
}

template <typename R>
using PKernel = void (*)(PArgs<R>);

template <typename R>
PKernel<R> pk_select(int ndir, bool tang, bool prim) {
  if (ndir == 3) {
    if (tang) return prim ? k_persist<3, true, true, R> : k_persist<3, true, false, R>;
    return prim ? k_persist<3, false, true, R> : k_persist<3, false, false, R>;
  }
  if (tang) return prim ? k_persist<2, true, true, R> : k_persist<2, true, false, R>;
  return prim ? k_persist<2, false, true, R> : k_persist<2, false, false, R>;
}

// Largest grid whose work-groups are all resident (cooperative launch).
template <typename R>
int persistent_grid(int ndir, bool tang, bool prim, int ntile, int* grid) {
  int dev = 0, ncu = 0, per = 0;
  if (hipGetDevice(&dev) != hipSuccess) return FUSS_EHIP;
  if (hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess) return FUSS_EHIP;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per, reinterpret_cast<const void*>(pk_select<R>(ndir, tang, prim)),
                                                   TILE_T, 0) != hipSuccess)
    return FUSS_EHIP;
  const int cap = per * ncu;
  if (cap < 1) return FUSS_EUNSUPPORTED;
  *grid = ntile < cap ? ntile : cap;
  return 0;
}

template <typename R>
int launch_persistent(fuss_ctx* c, const PArgs<R>& a, int ndir, bool tang, bool prim, int grid) {
  PArgs<R> args = a;
  void* kp[] = {&args};
  const hipError_t e = hipLaunchCooperativeKernel(reinterpret_cast<const void*>(pk_select<R>(ndir, tang, prim)),
                                                  dim3(grid), dim3(TILE_T), kp, 0, c->stream);
  return e == hipSuccess ? 0 : FUSS_EHIP;
}

template int persistent_grid<double>(int, bool, bool, int, int*);
template int persistent_grid<float>(int, bool, bool, int, int*);
template int launch_persistent<double>(fuss_ctx*, const PArgs<double>&, int, bool, bool, int);
template int launch_persistent<float>(fuss_ctx*, const PArgs<float>&, int, bool, bool, int);

}  // namespace fuss
