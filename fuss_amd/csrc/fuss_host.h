// fuss_host.h -- host-side context, level and buffer types of the FUSS
// explicit step library, and the launchers each kernel unit provides.
#pragma once

#include <vector>

#include "fuss_dev.h"
#include "fuss_refine_types.h"

namespace fuss {

template <typename T>
struct DBuf {
  T* p = nullptr;
  size_t n = 0;
  DBuf() = default;
  DBuf(const DBuf&) = delete;
  DBuf& operator=(const DBuf&) = delete;
  DBuf(DBuf&& o) noexcept : p(o.p), n(o.n) { o.p = nullptr; o.n = 0; }
  DBuf& operator=(DBuf&& o) noexcept {
    if (this != &o) { free(); p = o.p; n = o.n; o.p = nullptr; o.n = 0; }
    return *this;
  }
  ~DBuf() { free(); }
  int alloc(size_t count) {
    free();
    n = count;
    if (!count) return 0;
    if (hipMalloc(&p, sizeof(T) * count) != hipSuccess) { p = nullptr; n = 0; return FUSS_EHIP; }
    // the zero fill runs on the null stream, which the library's non-blocking
    // stream does not wait for: finish it before any kernel on that stream
    // writes the buffer (else the fill can land after, over the results)
    if (hipMemset(p, 0, sizeof(T) * count) != hipSuccess) return FUSS_EHIP;
    if (hipStreamSynchronize(nullptr) != hipSuccess) return FUSS_EHIP;
    return 0;
  }
  void free() {
    if (p) (void)hipFree(p);
    p = nullptr;
    n = 0;
  }
};

struct HostBlock {
  int n[3] = {0, 0, 0};
  std::vector<double> nodes, T, matID, qvol;
};

struct Level {
  int nb = 0;
  int kp = 2;                       // k ghost layers (1 in 2-D: k-face data only)
  std::vector<BlockInfo> hblk;
  std::vector<HostBlock> hb;        // host nodes (every level), fields (level 1)
  std::vector<Tile> htiles;         // one 64x4 cell tile per plane
  std::vector<Tile> hctiles;        // 64x4 columns of kc planes (3-D levels)
  std::vector<Tile> hc8tiles;       // 64x8 columns of kc planes (fused x-y-z IRS pass)
  int kc = KC_DEFAULT;              // planes per column tile (fuss_autotune)
  int xyz_ty = 4;                   // rows of the fused IRS pass's columns, 4 or 8 (fuss_autotune)
  int stage_ty = 4;                 // rows of the LDS-ring stage kernel's columns, 4 or 8 (fuss_autotune)
  int dt_ty = 4;                    // ... of the pass that wrote the dt minima (dt_next)
  int row_march = -1;               // orthogonal 3-D stage kernel: -1 context default, 0 LDS ring, 1 row march
  std::vector<Tile> hxytiles;       // 64x16 plane tiles of the x-y IRS pass
  std::vector<int> hrng[4];         // per tile list (tiles, ctiles, xytiles, c8tiles): first tile of each block
  DBuf<int> rng[4];
  int plist = 0;                    // tile list of the last step's residual partials
  int no_ycon = 1;                  // no connection on a j face (fused x-y IRS allowed)
  int no_zcon = 1;                  // no connection on a k face (with no_ycon: fused x-y-z IRS)
  int no_xcon = 1;                  // no connection on an i face
  // line-tridiagonal IRS: line tiles (x: 64 rows of a plane; y, z: 64x4
  // line origins) and the per-block elimination factors for tri_beta
  std::vector<Tile> htx, hty, htz;
  DBuf<Tile> ttx, tty, ttz;
  DBuf<double2> tri;
  DBuf<int> trioff;
  double tri_beta = -1.0;
  int ghost_read = 1;               // some BC entry reads layer-1 ghosts (101, 300, 302)
  int dt_next = 0;                  // dt and its per-tile minima hold the next step's (fused IRS pass)
  double dt_vnn = 0.0;              // ... computed with this vnn
  int all_umat = 0;                 // every block has a single material
  int all_uqv = 0;                  // every block has a uniform qvol
  int dims2 = 0;                    // every block has >= 2 cells in each direction
  // rectilinear levels: per-axis metric factors (RectAxes) that reproduce the
  // stored volumes and normal weights to rect_dev (relative) and sum 1/dl^2 exactly
  DBuf<double> rfac;
  int rect_w = 0;                   // every block rectilinear, factors within RECT_TOL
  int rect_dt = 0;                  // ... and sum 1/dl^2 bit-identical
  double rect_dev = -1.0;           // largest relative deviation found (-1: not checked)
  long npad = 0, ncell = 0, nnodes = 0;
  DBuf<BlockInfo> blk;
  DBuf<Tile> tiles, ctiles, xytiles, c8tiles;
  DBuf<double> nodes;
  DBuf<double> Tb[3];               // T^n and two stage buffers
  int cur = 0;                      // index of T^n in Tb
  DBuf<double> R, R2, dt, qvol, vol, invL2;
  DBuf<int> mat;
  DBuf<double> M, F[3], W[9];
  DBuf<BCRec> bc;
  DBuf<BCHot> bch;                  // compact wall records (k_bcflux)
  // connection (101) entries as structure-of-arrays records for k_connflux:
  // CONN_ND doubles and CONN_NI ints per entry, field-major (field f of
  // entry q at f * nconn + q)
  DBuf<double> cd;
  DBuf<int> ci;
  int nconn = 0;
  DBuf<int2> gmap;                  // (ghost, source) per entry (k_ghost)
  DBuf<int> fmap;
  DBuf<double> bcv, bcf;
  int nbound = 0;
  std::vector<int> hib;             // nbound x 15
  // chimera (102) donor lists (fuss_set_chimera, host): per listed entry
  // its BC index and donor counts, then 4 ints (b, i, j, k) and a volume
  // fraction per donor, entries in list order
  std::vector<int> hch_ent, hch_n1, hch_n2, hch_don;
  std::vector<double> hch_vf;
  // device: per chimera entry (ghost cell, first donor, layer-1 donors,
  // material index), the donors' arena indices and fractions
  DBuf<int4> chr;
  DBuf<int> chd;
  DBuf<double> chv;
  int nchim = 0;
  std::vector<double> hrb;          // nbound x 6
  std::vector<int> hser;
  int tang = 0, any_active = 0;
  double time = 0.0;                // host mirrors of the device clock
  int iter = 0;
  DBuf<double> clk;                 // [0] level time
  DBuf<int> cnt;                    // [0] iteration, [1] halt, [2] run steps, [3] chunk start, [4] run step limit
  // fuss_run's captured chunk of RUN_G steps, replayed by later calls with the
  // same step parameters and stop thresholds (dropped when tiles change)
  hipGraph_t run_graph = nullptr;
  hipGraphExec_t run_exec = nullptr;
  fuss_step_params run_p{};
  double run_thr = 0.0, run_tthr = 0.0;
  int run_seq[RUN_G + 1] = {};
  int run_g = 0;                    // steps in the captured chunk (<= RUN_G, multiple of 3)
  std::vector<const void*> run_bufs;  // device buffers the captured steps address (checked before a replay)
  DBuf<double> partials, pmin, red2;
  // FP32 explicit steps (FUSS_FP32 contexts): the per-cell fields in single
  // precision.  state_f: the float T buffers hold the level's current state
  // (the double ones are stale until state_to_f64); f_ready: the static
  // fields (volumes, weights, qvol) are converted.
  struct F32 {
    DBuf<float> Tb[3], R, R2, dt, qvol, vol, W[9], bcv, bcf;
  } f;
  int state_f = 0, f_ready = 0;
  int has_conn = 0;                 // some BC entry is a connection (101) or chimera (102) entry
  // persistent steps (fuss_persistent): barrier and control words, parity
  // halves of the tile minima and residual partials, per-step history
  DBuf<unsigned char> pk;
  size_t pk_hist = 0;               // steps of history the buffer holds
};

// The per-cell fields one precision's step reads and writes.
template <typename R>
struct FieldView {
  R* Tb[3];
  R *res, *res2, *dt, *qvol, *vol, *W[9], *bcv, *bcf;
  double* invL2;  // sum 1/dl^2 stays FP64 (the dt is formed in FP64 in either precision)
};

// One level of the correction multigrid (arrays in the inner precision).
struct CLevel {
  int ncell = 0, maxcells = 0;
  std::vector<CBlk> hblk;
  DBuf<CBlk> blk;
  DBuf<unsigned char> mem;
  void* d[3] = {nullptr, nullptr, nullptr};
  void *b = nullptr, *cx = nullptr, *cy = nullptr, *cz = nullptr, *D = nullptr, *s = nullptr, *s2 = nullptr;
  int cur = 0;
  // conjugate gradients on a coarsest level (refine coarse_cg > 0): r, z,
  // two direction buffers, A p, 1/diag; per-group partial sums (two of
  // r . z, one of p . q) and the last two r . z
  DBuf<unsigned char> cgmem;
  void* cg[6] = {nullptr, nullptr, nullptr, nullptr, nullptr, nullptr};
  DBuf<double> cgpart;
};

struct Refine {
  int ready = 0;
  fuss_refine_params p{};
  std::vector<CLevel> lv;
  DBuf<double> bcd, red;
  DBuf<double> bres;                // per-block monitor residual (independent cases)
  double* hbres = nullptr;          // pinned
  std::vector<int> conv_outer;      // first outer iteration at which each block met the threshold
  std::vector<double> res;          // last monitor residual per block
};

// Checkpointed forward and reverse sweep of the time-accurate step.
struct Adjoint {
  int ready = 0, prec = FUSS_FP64;
  int nsteps = 0, K = 0, nprobe = 0, forward_done = 0;
  std::vector<long> hprobe;
  DBuf<long> probe, coff;
  DBuf<double> ckpt;                // (nsteps / K + 1) states
  DBuf<double> seg;                 // the K states of one segment
  DBuf<double> stage[4];            // stage inputs of the recomputed step
  DBuf<unsigned char> lam;          // lam, lin, lTO in the adjoint precision
  DBuf<double> part, red, acc;      // partial sums; acc[0] dL/d(dt), acc[1] dL/dc
  DBuf<int> amin;
  std::vector<double> times;        // level time before each step
  std::vector<int> iters;
  DBuf<double> dtimes;              // the same, recorded on the device during the forward
  DBuf<int> diters;
  fuss_step_params p{};
};

}  // namespace fuss

struct fuss_ctx {
  int nlev = 0, nb = 0, precision = FUSS_FP64, ndir = 0;
  int dev = 0;                      // HIP device the context was created on
  std::vector<Level> lev;
  hipStream_t stream = nullptr;
  bool own_stream = false;
  int nmat = 0, Ti1 = 0, nT = 0;
  std::vector<double> tcp, trho, tkap, th;
  DBuf<double> dcp, drho, dkap, dh, dkap2, dpr4;
  DBuf<float> fcp, frho, fkap, fh, fkap2, fpr4;  // the tables in FP32 (FUSS_FP32 steps)
  int ftab_ok = 0;                  // ... converted from the current double tables
  struct Series { int periodic; std::vector<double> t, v; };
  std::vector<Series> series;
  DBuf<int> status;                 // code, b, i, j, k
  DBuf<double> scal;                // dtglobal, time, res_sq
  double* hscal = nullptr;          // pinned readback
  int setup_done = 0;
  int* hstatus = nullptr;
  Refine ref;                       // steady defect correction (fuss_refine)
  Adjoint adj;                      // discrete adjoint (fuss_adj_*)
  int adj_store_on = 0;             // enqueue_step copies each stage input to adj.stage[]
  int graphs_off = 0;               // fuss_run without graph capture (fuss_graphs)
  int persist = 0;                  // fuss_run as persistent launches where eligible (fuss_persistent)
  int fuse_small = 1;               // fused stages on 2-D levels without IRS and connections (fuss_fused)
  int stage_lds = 1;                // orthogonal 3-D stages: LDS-ring kernel (1) or row march (0)
  int rect_off = 0;                 // per-axis metric factors switched off (fuss_rect)
  double kscale = 1.0;              // conductivity table scale (fuss_set_kscale)
  DBuf<int> stage_i;                // staging for fuss_set_bc_values
  DBuf<double> stage_d;
  DBuf<double> ser_t, ser_v;        // time series on the device
  DBuf<int> ser_off, ser_n, ser_per;
  DBuf<double> run_hist;            // per-chunk residual / time history of fuss_run
  double* hrun = nullptr;           // pinned: run history readback
  int* hcnt = nullptr;              // pinned: level counters readback
  // per-kernel-class timing with HIP events on the context stream
  struct Prof {
    int enabled = 0;
    std::vector<hipEvent_t> pool;
    std::vector<int> cls;       // class of each recorded (start, end) pair
    size_t used = 0;
  } prof;
};

namespace fuss {

// Event pair around one launch of a kernel class (profiling mode only).
int prof_begin(fuss_ctx* c, int cls);
void prof_end(fuss_ctx* c, int id);

// Launchers (one per kernel family; defined in the stage_*.hip and irs_*.hip units).
template <typename R> void launch_stage2d(fuss_ctx* c, Level& L, const StageArgs<R>& a, bool tang, bool prim, bool irs);
template <typename R> void launch_stage3d(fuss_ctx* c, Level& L, const StageArgs<R>& a, bool tang, bool prim, bool irs);
template <typename R, int TY> void launch_stage3p_ring(fuss_ctx* c, Level& L, const StageArgs<R>& a, bool prim, bool irs);
template <typename R> void launch_stage_raw(fuss_ctx* c, Level& L, const StageArgs<R>& a, int ndir, bool tang);
template <typename R> void launch_newstate(fuss_ctx* c, Level& L, const StageArgs<R>& a, bool prim, bool irs);
template <typename R> void launch_irs_xyz(fuss_ctx* c, Level& L, const IrsArgs<R>& ia, bool prim);
template <typename R, int TY> void launch_irs_xyz_ty(fuss_ctx* c, Level& L, const IrsArgs<R>& ia, bool prim);
template <typename R> void launch_irs_z(fuss_ctx* c, Level& L, const IrsArgs<R>& ia, bool prim);
template <typename R> void launch_irs_line(fuss_ctx* c, Level& L, int ndir, int dir, const IrsArgs<R>& a, bool update, bool prim);
template <typename R> void launch_irs_xy(fuss_ctx* c, Level& L, int ndir, const IrsArgs<R>& ia, bool prim);
// Persistent steps (persist.hip): kernel arguments, resident grid size and
// the cooperative launch.
template <typename R>
struct PArgs {
  StageArgs<R> sa;       // level arrays; T/TO/Tout and the stage fields are set per stage
  DtArgs<R> da;
  const BCHot* hot;
  const int2* gmap;      // (ghost, source) per BC entry
  R* Tb[3];
  int cur0;              // index of T^n at the launch
  int nrk, jam;
  double rkc[4];
  int ntile;
  int time_accurate;
  int ghosts;            // write the layer-1 ghosts of each new state
  int compute_residual;
  int maxcount;
  double thr, tthr;
  double time0;
  int iter0;
  double* clk;           // level clock [0] time
  int* cnt;              // level counters [0] iteration
  double* scal;          // [0] dtglobal, [1] time, [2] Sum (T-TO)^2
  double* hres;          // per-step residual and time (or null)
  double* htime;
  double* pmin2;         // 2 x ntile: tile dt minima by step parity
  double* part2;         // 2 x ntile: residual partials by step parity
  unsigned* bar;         // [0] arrivals, [1] releases
  int* ctl;              // [0] abort, [1] status snapshot, [2] steps done, [3] final buffer index, [4] stopped
};
template <typename R>
void launch_stage_fused(fuss_ctx* c, Level& L, const StageArgs<R>& a, const DtArgs<R>& da, int ndir, bool tang,
                        bool prim, bool dt_here, bool ghosts);
void launch_fused_end(fuss_ctx* c, Level& L, int n, bool dt_here, int time_accurate);
template <typename R> int persistent_grid(int ndir, bool tang, bool prim, int ntile, int* grid);
template <typename R> int launch_persistent(fuss_ctx* c, const PArgs<R>& a, int ndir, bool tang, bool prim, int grid);
template <typename R> int launch_irs_tridiag(fuss_ctx* c, Level& L, IrsArgs<R> ia, R* rin, R* rout, bool prim, int* npart);

}  // namespace fuss

