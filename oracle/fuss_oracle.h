/*
 * fuss_oracle.h -- CPU restatement (FP64) of the FUSS explicit solver step.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker for the HIP path:
 * only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may
 * load it.  The product (fuss_amd, libfuss_gpu.so) never links or calls it.
 *
 * Every function restates one routine of the reference (open-hydra/FUSS,
 * Fortran 2008); the .c file cites file:line for each.  Arithmetic is kept
 * in the reference's source order and the file is compiled with
 * -ffp-contract=off so no fused multiply-adds are introduced.
 *
 * Indices follow the reference: blocks, cells and faces are 1-based, cell
 * arrays carry gc=2 ghost layers on every side, faces are numbered
 * 1=i-min, 2=i-max, 3=j-min, 4=j-max, 5=k-min, 6=k-max.
 */
#ifndef FUSS_ORACLE_H
#define FUSS_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

typedef struct ob_ctx ob_ctx;

/* Integer layout of one BC entry (15 ints), as read from INPUT/bc.txt:
 * b, i, j, k, f, type, bs, is, js, ks, fs, d11, d12, d21, d22. */
#define OB_BC_NI 15
/* Real layout of one BC entry (6 doubles): qw, Tw, hconv, qrad, Tref, eps. */
#define OB_BC_NR 6

/* Status codes. */
#define OB_OK 0
#define OB_ERR_ARG 1
#define OB_ERR_STATE 2      /* NaN or T<0 after an update (Check_And_Fix_State) */
#define OB_ERR_TABLE 3      /* temperature or enthalpy outside the property table */
#define OB_ERR_METRIC 4     /* det(M)=0 or vol<=0 */
#define OB_ERR_UNSUPPORTED 5

ob_ctx *ob_create(int nlevels, int nb);
void ob_destroy(ob_ctx *c);

/* Property tables, one row per integer kelvin starting at Ti1, nT rows,
 * nmat materials, each array nmat*nT with the temperature index fastest. */
int ob_set_table(ob_ctx *c, int nmat, int Ti1, int nT, const double *cp,
                 const double *rho, const double *kap, const double *h);

/* Finest-level block b (1-based): nodes (ni+1)(nj+1)(nk+1) x 3 with the
 * coordinate fastest then i, j, k; T, matID, qvol interior ni*nj*nk with i
 * fastest (qvol may be NULL = 0). */
int ob_set_block(ob_ctx *c, int b, int ni, int nj, int nk, const double *nodes,
                 const double *T, const double *matID, const double *qvol);

/* Time series used by time-dependent 301/302 values. Returns its id >= 0. */
int ob_add_series(ob_ctx *c, int periodic, int n, const double *t, const double *v);

/* BC list of one level (1-based level), entries grouped by (block, face)
 * ascending as in the reference's bc files.  series may be NULL. */
int ob_set_bc(ob_ctx *c, int level, int nbound, const int *ib, const double *rb,
              const int *series);

/* Numerics: n_rk in {1,2,3} (reference RK), 4 = Jameson four-stage; prim=1 primitive, 0 conservative; irs_beta>0
 * enables implicit residual smoothing. */
/* chimera donor lists of the level's 102 entries: entry[q] (0-based BC index),
 * n1[q] + n2[q] donors (b, i, j, k; 1-based) with volume fractions, in entry
 * order (IO_BC.f90:218-229) */
int ob_set_chimera(ob_ctx *c, int level, int n, const int *entry, const int *n1, const int *n2,
                   const int *donors, const double *vf);
int ob_set_params(ob_ctx *c, int n_rk, double vnn, int rampa_iter,
                  int time_accurate, int prim, double irs_beta);

/* IRS flavour: 0 the reference's two Jacobi sweeps per direction, 1 the
 * exact line-tridiagonal solve with the same ghosts (no reference
 * counterpart). */
int ob_set_irs_mode(ob_ctx *c, int mode);

/* Coarse levels, metrics, matIDg, initial ghost fill and restriction, as
 * FUSS_setup does.  newrun_time: initial time (0 transient, -1 steady). */
int ob_setup(ob_ctx *c, double initial_time);

/* One Explicit_Step on `level`; returns Sum (T-TO)^2 in *res_sq. */
int ob_step(ob_ctx *c, int level, double *res_sq);

int ob_restrict(ob_ctx *c, int fine_level);
int ob_prolong(ob_ctx *c, int fine_level);

/* Driver: FUSS_solve/FUSS_postprocess control for steady or transient
 * runs with coarse-to-fine multigrid sequencing.  itermax[m-1] is the
 * iteration limit of level m (level 1 gets iter_threshold).  Writes one
 * history row per iteration (iter_general, time, residual) up to hist_cap
 * rows; probes (level-1 cells, 4 ints each b,i,j,k) are sampled after every
 * level-1 iteration into probe_out[row*nprobe + p].  Returns status;
 * *n_iter receives the number of iterations taken. */
int ob_run(ob_ctx *c, const int *itermax, double res_threshold,
           double time_threshold, int max_iters, int hist_cap, int *n_iter,
           int *hist_iter, double *hist_time, double *hist_res, int nprobe,
           const int *probes, double *probe_out);

/* Field access (level, block 1-based).  name: "T","TO","R","dt","vol",
 * "qvol","matID" -> interior ni*nj*nk (i fastest); "Tpad" -> padded. */
int ob_get(ob_ctx *c, int level, int b, const char *name, double *out);
int ob_set(ob_ctx *c, int level, int b, const char *name, const double *in);
int ob_dims(ob_ctx *c, int level, int b, int *dims);
int ob_ndir(ob_ctx *c);
int ob_threads(void);
double ob_time(ob_ctx *c, int level);
double ob_dtglobal(ob_ctx *c, int level);
int ob_last_error(ob_ctx *c, int *bijk);

/* Metric outputs of block b at a level: M (9/cell, padded, c(3,3) column
 * major), dl (3/cell padded), vol (padded), face N (3/face) and A of
 * direction d (1..3, sizes of Allocate_Block).  Any pointer may be NULL. */
int ob_get_metrics(ob_ctx *c, int level, int b, double *M, double *dl,
                   double *vol, int d, double *fN, double *fA);
/* Mg(1) and matIDg of BC entry n (1-based) of a level. */
/* Wall temperature and heat flux of every BC entry of `level` at the
 * current state (Extract_Properties_From_BC, IO_Wall.f90:131-239; entries
 * that are not walls 301-305 get 0), in BC-list order. */
int ob_wall(ob_ctx *c, int level, double *Tw, double *qw);
int ob_get_bc_ghost(ob_ctx *c, int level, int n, double *Mg1, double *matIDg,
                    double *Tg);

#ifdef __cplusplus
}
#endif
#endif
