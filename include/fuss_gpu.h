/*
 * fuss_gpu.h -- C-ABI of the MI355X (gfx950) explicit solver step of FUSS.
 *
 * Drop-in boundary for the reference's hot path (open-hydra/FUSS, Fortran):
 * its Fortran driver keeps FUSS_setup / FUSS_solve / FUSS_postprocess, the INI
 * configuration and the BC / material files, and calls these entry points
 * through iso_c_binding (fuss_amd/fortran/fuss_gpu_m.f90) instead of the
 * OpenMP loops of Explicit_Step.  Plain C types only: int, double, pointers
 * and sizes.  Every entry returns 0 on success or a FUSS_E* status.
 *
 * Index conventions are the reference's: blocks, cells and levels are
 * 1-based; faces are 1=i-min, 2=i-max, 3=j-min, 4=j-max, 5=k-min, 6=k-max;
 * cell arrays passed in or out hold the interior ni*nj*nk cells with i
 * fastest (Fortran T(1:ni,1:nj,1:nk)); node arrays hold (ni+1)(nj+1)(nk+1)
 * points x,y,z with the coordinate fastest (Fortran node(0:ni,0:nj,0:nk)%c).
 */
#ifndef FUSS_GPU_H
#define FUSS_GPU_H

#ifdef __cplusplus
extern "C" {
#endif

typedef struct fuss_ctx fuss_ctx;

/* status codes */
#define FUSS_OK 0
#define FUSS_EARG 1          /* invalid argument or call order */
#define FUSS_ESTATE 2        /* NaN or T<0 after an update: Check_And_Fix_State */
#define FUSS_ETABLE 3        /* T or h outside the property table */
#define FUSS_EMETRIC 4       /* det(M)=0 or vol<=0 in the metrics */
#define FUSS_EUNSUPPORTED 5  /* e.g. chimera (102) entries */
#define FUSS_EHIP 6          /* HIP runtime error */
#define FUSS_EBC 7           /* BC list inconsistent with the blocks */

/* precision modes of a context */
#define FUSS_FP64 0          /* plain FP64 path: parity with the reference */
#define FUSS_FP32 1          /* FP32 explicit steps: T, fluxes, RK and IRS in single
                                precision; the local dt is formed in FP64 and both
                                reductions (dt min, Sum (T-TO)^2) and the clock stay
                                FP64.  fuss_refine's monitor and the adjoint's
                                forward run FP64 steps; hooks and chimera entries
                                are FP64 only (FUSS_EUNSUPPORTED). */

/* integration variables (Mod_Newstate.f90:191-205) */
#define FUSS_CONS 0
#define FUSS_PRIM 1

/* IRS flavours */
#define FUSS_IRS_JACOBI2 0   /* reference: 2 Jacobi sweeps per direction */
#define FUSS_IRS_TRIDIAG 1   /* exact line-tridiagonal solve per direction */

/* Numerics of one step.  Replaces the obj_time_scheme / obj_irs globals that
 * Explicit_Step reads (src/lib/config/Config_Types_m.f90, Register_Numerics.f90:30-48). */
typedef struct {
  int n_rk;            /* 1 euler, 2 RK2, 3 RK3 (Lib_RK.f90:8-11), 4 = four-stage extension */
  double vnn;          /* VNN number (Mod_dt.f90:262-288) */
  int rampa_iter;      /* vnn-rise-threshold */
  int time_accurate;   /* global dt (Set_Global_dt) when nonzero */
  int integration;     /* FUSS_CONS or FUSS_PRIM */
  double irs_beta;     /* IRS enabled when > 0 */
  int irs_mode;        /* FUSS_IRS_JACOBI2 or FUSS_IRS_TRIDIAG */
  int compute_residual;/* nonzero: return Sum (T-TO)^2 (Mod_Explicit.f90:65-78) */
} fuss_step_params;

/* Outputs of one step. */
typedef struct {
  double res_sq;       /* Sum over the level of (T^{n+1}-T^n)^2 */
  double dtglobal;     /* global min of the local dt (1e5 cap), Mod_dt.f90:228-258 */
  double time;         /* domain(level)%time after the step */
  int iter;            /* domain(level)%iter after the step */
  int status;          /* FUSS_OK or FUSS_ESTATE / FUSS_ETABLE */
  int bijk[4];         /* first failing cell (block, i, j, k) */
} fuss_step_result;

/* --- construction (replaces FUSS_setup's Setup_Data_Structure, Setup_BC,
 *     Setup_Multigrid, Setup_Metrics and first Fill_Ghost_Cell:
 *     src/lib/driver/Wrap_Setup.f90:10-165) ------------------------------ */

/* New context for nlevels multigrid levels of nblocks blocks. */
int fuss_create(int nlevels, int nblocks, int precision, fuss_ctx **out);

/* Property tables, Load_Table (src/lib/io/Load_Table.f90:10-75): nmat
 * materials x nT rows of 1 K starting at Ti1 >= 0; arrays nmat*nT, T fastest. */
int fuss_set_table(fuss_ctx *ctx, int nmat, int Ti1, int nT, const double *cp,
                   const double *rho, const double *kappa, const double *h);

/* Finest-level block b: dims, nodes, initial T, matID (material number as a
 * real, FUSS_block_type%matID), qvol (may be NULL).
 * Replaces Setup_Data_Structure (src/lib/driver/Mod_Allocate_Data.f90:7-60). */
int fuss_set_block(fuss_ctx *ctx, int b, int ni, int nj, int nk, const double *nodes,
                   const double *T, const double *matID, const double *qvol);

/* Time series of time-dependent 301/302 values, time_series_type
 * (src/lib/base/Series_Data_m.f90:20-86).  *id receives its index. */
int fuss_add_series(fuss_ctx *ctx, int periodic, int n, const double *t, const double *v,
                    int *id);

/* BC list of a level as read by Read_BCfile (src/lib/io/IO_BC.f90:121-245),
 * entries grouped by (block, face) ascending.  ib: nbound x 15 ints
 * (b,i,j,k,f,type,bs,is,js,ks,fs,d11,d12,d21,d22); rb: nbound x 6 doubles
 * (qw,Tw,hconv,qrad,Tref,eps); series: nbound ids or -1 (may be NULL). */
int fuss_set_bc(fuss_ctx *ctx, int level, int nbound, const int *ib, const double *rb,
                const int *series);

/* Chimera (102) donor lists of a level's BC entries (IO_BC.f90:218-229,
 * read into bc(i)%ni, %donorID, %volume_fraction): n entries, entry[q] the
 * 0-based BC index of a 102 entry, n1[q] donors for its first ghost layer and
 * n2[q] for the second, then per donor 4 ints (block, i, j, k; 1-based) in
 * `donors` and its volume fraction in `vf`, entries in list order.  Call after
 * fuss_set_bc and before fuss_setup; finest level only (FUSS_EUNSUPPORTED for
 * 102 entries on coarser levels).  The ghost fill forms the first layer from
 * the donors' enthalpies (Ghost_Chimera, Lib_Ghost.f90:192-245); the
 * boundary flux is the connection flux (Mod_BC_Fluxes.f90:45). */
int fuss_set_chimera(fuss_ctx *ctx, int level, int n, const int *entry, const int *n1, const int *n2,
                     const int *donors, const double *vf);

/* Coarse grids, metrics, ghost metrics, initial ghost fill and restriction
 * of the initial state to coarse levels.  initial_time: 0 for time-accurate
 * runs, -1 for steady runs (Wrap_Setup.f90:123-144). */
int fuss_setup(fuss_ctx *ctx, double initial_time);

/* --- hot path -------------------------------------------------------- */

/* One Explicit_Step on `level` (src/lib/numerics/time/explicit/Mod_Explicit.f90:10-111):
 * Compute_dt + min, Copy_State, n_rk x (Fill_Ghost_Cell, Fluxes, BC_Fluxes,
 * RK_Newstate [with Residual_Smoothing]), Compute_Residual. */
int fuss_step(fuss_ctx *ctx, int level, const fuss_step_params *p, fuss_step_result *r);

/* The External_Function slot of Explicit_Step (Mod_Explicit.f90:24,53-59): called
 * once per stage and block between BC_Fluxes and RK_Newstate with the block's
 * residual R -- the sum of its face fluxes and BC contributions as BC_Fluxes
 * leaves it (rho cp V dT/dt = -R + qvol V) -- and the stage's input T, both
 * ni*nj*nk doubles on the host in i-fastest order; R may be modified in place
 * (source terms).  block is 1-based, irk = 1..n_rk. */
typedef void (*fuss_stage_fn)(void *user, int level, int irk, int block, int ni, int nj, int nk,
                              const double *T, double *R);
/* fuss_step with the hook: the stage runs unfused (raw R to HBM, a host round
 * trip per stage and block, then the newstate pass), so it is for coupling,
 * not speed.  Not capturable in a graph. */
int fuss_step_hook(fuss_ctx *ctx, int level, const fuss_step_params *p, fuss_stage_fn fn,
                   void *user, fuss_step_result *r);

/* Up to max_iters Explicit_Steps on `level`, stopping after the first step
 * whose residual sqrt(res_sq) <= res_threshold, whose time >= time_threshold,
 * or when iter_from_call reaches itermax (the endsim/endmg tests of
 * Mod_Explicit.f90:81-90).  hist_res/hist_time (may be NULL) receive one row
 * per step; *n_done the number of steps taken. */
int fuss_run(fuss_ctx *ctx, int level, const fuss_step_params *p, int max_iters,
             int itermax_left, double res_threshold, double time_threshold, int *n_done,
             double *hist_res, double *hist_time, fuss_step_result *last);

/* fuss_run replays its steps from a captured HIP graph (chunks of 30 steps,
 * or of the request rounded up to a multiple of 3; stop tests evaluated on
 * the device, so it stops at the same iteration as a step-by-step loop)
 * unless profiling is on, fewer than 6 steps are asked for, or graphs are
 * switched off here (enable = 0).  Default: on.  No reference counterpart
 * (the reference's step loop is Mod_Explicit.f90:10-111 called from
 * Wrap_Solve.f90:24). */
int fuss_graphs(fuss_ctx *ctx, int enable);

/* With enable = 1, fuss_run runs all its steps in one persistent cooperative
 * launch on levels without IRS and without connection (101/102) entries:
 * the work-groups step through dt, the RK stages (fused fluxes, BC values,
 * update and mirror ghosts) and the stop tests with grid-wide barriers, for
 * latency-bound levels where per-step launches dominate.  Same results,
 * stop iteration and history as the other fuss_run modes.  Other levels
 * keep the graph/loop path.  Default: off.  No reference counterpart
 * (replaces the per-step OpenMP region of Mod_Explicit.f90:48-62). */
int fuss_persistent(fuss_ctx *ctx, int enable);

/* On 2-D levels without IRS and without connection entries, each RK stage
 * is one fused launch: the stage evaluates its cells' BC values in place
 * (no surface pass), writes the layer-1 mirror ghosts of the new state (no
 * ghost pass before the next stage), and on steady steps the first stage
 * computes the local dt; one kernel closes the step (residual sum, dt
 * minimum, clock).  Same arithmetic and order as the split passes (which
 * enable = 0 restores).  Default: on.  No reference counterpart (the
 * passes are Mod_Explicit.f90:48-62's). */
int fuss_fused(fuss_ctx *ctx, int enable);

/* Rectilinear levels (3-D, every node of every block on its coordinate
 * planes x = x_i, y = y_j, z = z_k, no tangential weights): the 3-D stage
 * pass forms cell volumes and normal face weights as products of per-axis
 * factors instead of streaming the stored arrays, and the fused next-step dt
 * forms sum 1/dl^2 from them.  No reference counterpart (the reference reads
 * its 3-D metric arrays, Lib_Diffusive.f90:34-40, Mod_dt.f90:26-84): a storage
 * choice of this library.  fuss_setup checks the factors against the stored
 * metrics on the device and uses them only when every volume and weight
 * agrees to 1e-14 relative (sum 1/dl^2: bit for bit).  Default: on;
 * enable = 0 streams the stored arrays. */
int fuss_rect(fuss_ctx *ctx, int enable);
/* Whether `level` uses the factors for weights / dt, and the largest relative
 * deviation found by the setup check (-1: level not rectilinear). */
int fuss_rect_info(fuss_ctx *ctx, int level, int *weights, int *dt, double *max_rel_dev);

/* Tile autotuner: times `steps` Explicit_Steps with each candidate of the
 * 3-D stage pass -- column height 4, 8, 16, 32 or 64 planes x LDS-ring or row-march
 * kernel (row march: orthogonal levels only), then the column heights with
 * 8-row columns for the LDS-ring stage kernel and the fused x-y-z IRS pass
 * (where they run) -- on the level's
 * current state, restores the state and clock, and keeps the fastest for this level.  The
 * per-cell arithmetic is the same for every candidate; only the grouping of
 * the residual's partial sums changes.  2-D levels: nothing to tune (kc = 0). */
int fuss_autotune(fuss_ctx *ctx, int level, const fuss_step_params *p, int steps, int *kc,
                  int *row_march, double *ms_per_step);
/* Column tiles of a 3-D level without timing (no reference counterpart: the
 * reference has no tiles): kc = 4, 8, 16, 32 or 64 planes; row_march = -1
 * (context default), 0 (LDS ring) or 1 (row march); stage_rows = 4 or 8 rows
 * per column of the LDS-ring stage kernel (orthogonal blocks), or
 * FUSS_STAGE_REGCOL: the register-column stage kernel (one wave per 64x4
 * column, no LDS ring and no barriers; rectilinear levels with kc = 8, 16 or
 * 32 -- other levels run the 4-row LDS ring); irs_rows = 4 or 8 rows per
 * column of the fused x-y-z IRS pass (levels without j/k connections, Jacobi
 * IRS).  fuss_autotune also chooses the rows and kernels; fuss_tiles reports
 * the level's current choice (0s on 2-D levels).  Results do not depend on
 * the tiles except for the grouping of the residual's partial sums. */
#define FUSS_STAGE_REGCOL 1
int fuss_set_tiles(fuss_ctx *ctx, int level, int kc, int row_march, int stage_rows, int irs_rows);
int fuss_tiles(fuss_ctx *ctx, int level, int *kc, int *row_march, int *stage_rows, int *irs_rows);

/* Restriction (Mod_Multigrid.f90:59-77): coarse T, qvol, matID from level
 * fine_level to fine_level+1. */
int fuss_restrict(fuss_ctx *ctx, int fine_level);
/* Prolongation (Mod_Multigrid.f90:80-95): fine T of fine_level replaced by
 * the interpolation of level fine_level+1. */
int fuss_prolong(fuss_ctx *ctx, int fine_level);

/* --- steady state by FP64 defect correction (no reference counterpart) -----
 * The reference reaches a steady state only by marching Explicit_Step until
 * its residual ||T^{n+1}-T^n||_2 (Mod_Explicit.f90:65-90) drops below
 * res-threshold.  fuss_refine reaches the same stopping test faster: each
 * outer iteration takes one FP64 Explicit_Step (the monitor: its residual is
 * the reference's), evaluates the FP64 defect -R + qvol V of the finest
 * level, solves the linearised correction equation with a multigrid V-cycle
 * whose multistage smoother sweeps run in `precision`, and adds the
 * correction to T in FP64.  Finest level only; blocks must not be joined by
 * connections (FUSS_EUNSUPPORTED otherwise). */
typedef struct {
  int precision;       /* FUSS_FP32 or FUSS_FP64: inner sweeps */
  int n_stage;         /* stages of the multistage smoother (1..8) */
  double lam_lo, lam_hi; /* interval of the Jacobi-scaled spectrum the stages damp */
  double irs_beta;     /* IRS (two Jacobi sweeps per direction) on each stage, 0 = off */
  int nu1, nu2;        /* pre / post smoothing iterations */
  int vcycles;         /* V-cycles per outer iteration */
  int max_levels;      /* correction levels, 0 = coarsen to <= 2 cells per direction */
  int ncoarse;         /* smoothing iterations on a coarsest level too large for the direct solve */
  int max_outer;       /* outer iterations (monitor steps - 1) */
  double res_threshold;/* stop when the monitor residual sqrt(res_sq) <= this */
  int per_block;       /* nonzero: blocks are independent cases, each must meet res_threshold
                          with its own residual; 0: one residual over all blocks */
  int coarse_cg;       /* > 0: this many Jacobi-preconditioned conjugate-gradient iterations
                          solve a coarsest level too large for the direct solve (instead of
                          ncoarse smoothing iterations) */
} fuss_refine_params;

typedef struct {
  int outer;           /* monitor steps taken - 1 (= corrections applied) */
  int vcycles;         /* V-cycles run */
  int levels;          /* correction levels */
  int converged;       /* 1 when the monitor residual reached res_threshold */
  double res;          /* last monitor residual sqrt(res_sq) */
  double defect;       /* ||-R + qvol V||_2 (W) before the last correction */
} fuss_refine_result;

/* Per block: the first outer iteration at which its own monitor residual met
 * res_threshold (-1: never) and its last monitor residual (nblocks each). */
int fuss_refine_blocks(fuss_ctx *ctx, int *conv_outer, double *res);
/* Builds the correction hierarchy (allocates device memory). */
int fuss_refine_setup(fuss_ctx *ctx, const fuss_refine_params *rp);
/* Runs the outer iteration; `monitor` are the Explicit_Step numerics of the
 * monitor step (steady: time_accurate = 0).  hist_res (may be NULL) receives
 * the monitor residual of each outer iteration (max_outer + 1 entries). */
int fuss_refine(fuss_ctx *ctx, const fuss_step_params *monitor, fuss_refine_result *out,
                double *hist_res);

/* --- discrete adjoint of the time-accurate step (no reference counterpart) ---
 * Parameters of the forward model: wall values of BC entries (the reference
 * keeps them in bc%qw / bc%Tw ..., set once by Read_BCfile, IO_BC.f90:121-245)
 * and a scale of the conductivity table (Load_Table, Load_Table.f90:61-71). */
/* kappa table := scale x the table given to fuss_set_table. */
int fuss_set_kscale(fuss_ctx *ctx, double scale);
/* Sets column `column` of rb (0 qw, 1 Tw, 2 hconv, 4 Tref, 5 eps) for the
 * 0-based entries[n] of a level's BC list. */
int fuss_set_bc_values(fuss_ctx *ctx, int level, int n, const int *entries, int column,
                       const double *values);
/* Reverse-mode sweep over nsteps time-accurate Explicit_Steps of level 1 with
 * a checkpoint every ckpt_every steps; probe_bijk[4 nprobe] (1-based b,i,j,k)
 * are the cells whose T after every step the loss may depend on.  precision:
 * FUSS_FP64 or FUSS_FP32 adjoint fields (the forward recompute is FP64).
 * Scope: primitive variables, no IRS, orthogonal blocks, walls 0/301/302/303,
 * no connections (FUSS_EUNSUPPORTED otherwise). */
int fuss_adj_setup(fuss_ctx *ctx, int nsteps, int ckpt_every, int nprobe, const int *probe_bijk,
                   int precision);
/* nsteps steps from the current state, storing checkpoints; probe_hist
 * (device, nsteps x nprobe, may be NULL) receives T at the probes after each
 * step.  Leaves T^nsteps as the level state. */
int fuss_adj_forward(fuss_ctx *ctx, const fuss_step_params *p, double *probe_hist);
/* Given dL/d(probe_hist) (device, nsteps x nprobe), writes dL/dqw per BC entry
 * of level 1 (device, nbound; 301 entries without a time series), dL/dscale
 * (device, 1) and dL/dT^0 (device, interior cells block after block, i
 * fastest; may be NULL).  The level keeps T^nsteps. */
int fuss_adj_backward(fuss_ctx *ctx, const double *probe_adj, double *grad_qw, double *grad_kscale,
                      double *grad_T0);

/* --- state access (output cadence, restart, probes) ------------------- */
/* Wall output (Extract_Properties_From_BC, src/lib/io/IO_Wall.f90:131-239):
 * wall temperature and heat flux of each BC entry of `level` at the current
 * state, in BC-list order (nbound values each; 0 for entries that are not
 * walls 301-305).  Mirrors the level's ghosts from the current T first. */
int fuss_wall(fuss_ctx *ctx, int level, double *Tw, double *qw);

/* Largest BC list fuss_wall_noalloc stages (entries per level). */
#define FUSS_WALL_NOALLOC_MAX 262144
/* fuss_wall without any allocation: the device staging of Tw and qw is a
 * static per-GPU array of 2 x FUSS_WALL_NOALLOC_MAX doubles (4 MiB, held by
 * every GPU the library runs on until the process ends) instead of a buffer
 * allocated per call, so no hipMalloc / hipFree or allocating call is made.
 * No counterpart in the reference: for callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid it).  Levels with more than FUSS_WALL_NOALLOC_MAX entries are
 * rejected (FUSS_EARG).  Every context on a GPU shares that GPU's array:
 * calls that overlap on one GPU race. */
int fuss_wall_noalloc(fuss_ctx *ctx, int level, double *Tw, double *qw);
int fuss_get_T(fuss_ctx *ctx, int level, int b, double *T);
int fuss_set_T(fuss_ctx *ctx, int level, int b, const double *T);
/* Probes (src/lib/io/IO_Probes.f90:124-166): T of n level-1 cells bijk[4n]. */
int fuss_probe(fuss_ctx *ctx, int n, const int *bijk, double *out);
/* Level time / iteration counters (domain(level)%time, %iter). */
int fuss_get_clock(fuss_ctx *ctx, int level, double *time, int *iter);
int fuss_set_clock(fuss_ctx *ctx, int level, double time, int iter);

/* --- diagnostics / introspection ----------------------------------------- */
/* Number of dimensions from Check_Mesh_Type (Lib_Metrics.f90:13-66). */
int fuss_ndir(fuss_ctx *ctx, int *ndir);
/* Interior field copy of a level/block: "T","TO","R","dt","vol","qvol". */
int fuss_get_field(fuss_ctx *ctx, int level, int b, const char *name, double *out);
/* Cell metric tensor M (9 per interior cell, c(3,3) column-major), cell
 * volumes, and face normal/area of direction d (1..3) of a block, interior
 * ranges of Allocate_Block (Mod_Allocate_Data.f90:78-84).  NULL skips. */
int fuss_get_metrics(fuss_ctx *ctx, int level, int b, double *M, double *vol, int d,
                     double *N, double *A);
/* The HIP stream the context launches on (hipStream_t as void*), and a
 * caller-provided stream to use instead (borrowed, not destroyed). */
int fuss_get_stream(fuss_ctx *ctx, void **stream);
int fuss_set_stream(fuss_ctx *ctx, void *stream);
/* Per-kernel timing (no reference counterpart: replaces the reference's
 * Cpu_Time bracket, src/lib/driver/Wrap_Setup.f90:164, with HIP events on the
 * context stream).  fuss_profile(ctx, 1) starts recording an event pair
 * around every launch; fuss_profile_read synchronizes the stream and returns
 * the summed milliseconds and launch count per kernel class FUSS_K_*. */
#define FUSS_K_DT 0          /* local dt + global min */
#define FUSS_K_GHOST 1       /* ghost fill */
#define FUSS_K_STAGE 2       /* fused fluxes + BC fluxes + newstate */
#define FUSS_K_IRS 3         /* IRS line pass */
#define FUSS_K_IRS_UPDATE 4  /* last IRS line pass fused with the RK update */
#define FUSS_K_REDUCE 5      /* dt minimum and residual sum */
#define FUSS_K_BC 6          /* boundary-face fluxes (surface pass) */
#define FUSS_K_STAGE_DT 7    /* first stage computing the step's local dt (steady, rectilinear) */
#define FUSS_K_NCLASS 8
int fuss_profile(fuss_ctx *ctx, int enable);
int fuss_profile_read(fuss_ctx *ctx, int nclass, double *ms, int *count);
/* Device memory held by the context, bytes. */
int fuss_memory(fuss_ctx *ctx, double *bytes);
const char *fuss_strerror(int code);
void fuss_destroy(fuss_ctx *ctx);

#ifdef __cplusplus
}
#endif
#endif
