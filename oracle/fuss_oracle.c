/*
 * fuss_oracle.c -- CPU restatement (FP64) of the FUSS explicit solver step.
 *
 * TEST INFRASTRUCTURE ONLY (the checker of the HIP path; see fuss_oracle.h).
 *
 * Reference: open-hydra/FUSS, Fortran 2008.  Each routine below names the
 * reference file:line it restates.  Arithmetic follows the reference's source
 * order (Fortran evaluates a-b+c-d as ((a-b)+c)-d; matmul(G,M) as
 * (G1*M1x+G2*M2x)+G3*M3x; dot_product left to right); compile with
 * -ffp-contract=off.
 *
 * Pinning (see tests/test_oracle_golden.py): whole-trajectory residual
 * history of test/numerics-features/residual_convergence (MG + IRS +
 * connection), steady fields of test/steady-state/ (cons, prim, 301, 302,
 * 303, 304, qvol), transient multi-block multi-material fields of
 * test/transient/multimat_Plate, probe histories of
 * test/transient/oscillating_Rod (periodic time series), plus closed forms
 * (uniform-grid metrics, table interpolation, RK coefficients, IRS operator,
 * prolongation weights, restriction volume weighting).
 */
#ifdef _OPENMP
#include <omp.h>
#endif
#include "fuss_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

#define GC 2

/* src/lib/base/Parameters_m.f90:11-16 -- guide(6,3) face offsets. */
static const int GUIDE[6][3] = {{1, 0, 0}, {-1, 0, 0}, {0, 1, 0},
                                {0, -1, 0}, {0, 0, 1}, {0, 0, -1}};
/* src/lib/base/Parameters_m.f90:10 */
static const double SIGMA_SB = 5.67e-8;
/* src/lib/numerics/time/explicit/Lib_RK.f90:8-11 -- RKcoeff(n_rk, irk). */
static const double RKC[3][3] = {{1.0, 0.0, 0.0},
                                 {1.0, 0.50, 0.0},
                                 {1.0, 0.25, 2.0 / 3.0}};
/* n_rk = 4 (no counterpart in the reference, which stops at RK3): the
 * four-stage scheme of Jameson, Schmidt and Turkel (AIAA 81-1259),
 * new = old + a_k * residual(state), a = 1/4, 1/3, 1/2, 1.  For a linear
 * autonomous problem its amplification polynomial is the classical RK4 one,
 * 1 + z + z^2/2 + z^3/6 + z^4/24. */
static const double RK4A[4] = {0.25, 1.0 / 3.0, 0.5, 1.0};

typedef struct {
  int n[3];
  long s1, s2;      /* strides of padded cell arrays */
  long npad;
  double *T, *TO, *R, *RS1, *RS2, *qvol, *matID, *vol;
  double *M;        /* 9 per padded cell, c(r,c) at [(r-1)+3*(c-1)] */
  double *dl;       /* 3 per padded cell */
  double *dt;       /* interior ni*nj*nk */
  double *node;     /* (ni+1)(nj+1)(nk+1) x 3 */
  double *fN[3], *fA[3];
} ob_blk;

typedef struct {
  int i, j, k, b, f, type;
  int bs, is, js, ks, fs, d11, d12, d21, d22;
  double qw, Tw, hconv, qrad, Tref, eps;
  double Tg[6], matIDg;
  double Mg[2][9];
  int series;
  /* chimera (102): nch donors (b,i,j,k; 1-based) with volume fractions, the
   * first nch1 for ghost layer 1 (IO_BC.f90:218-229) */
  int nch1, nch;
  int *cdon;
  double *cvf;
} ob_bc;

typedef struct {
  int periodic, n;
  double *t, *v;
} ob_series;

typedef struct {
  int nb;
  ob_blk *blk;
  int nbound;
  ob_bc *bc;
  int *n_bf; /* nb x 6 */
  double time, dtglobal;
  int iter, itermax;
} ob_dom;

struct ob_ctx {
  int nlev, nb;
  ob_dom *dom;
  int nmat, Ti1, Ti2, nT;
  double *cp, *rho, *kap, *h;
  int ndir;
  int n_rk, rampa, time_accurate, prim, irs;
  int irs_mode;   /* 0: two Jacobi sweeps (reference), 1: exact line-tridiagonal solve */
  double vnn, beta;
  int nseries;
  ob_series *series;
  int err, eb, ei, ej, ek;
  int setup_done;
};

/* ------------------------------------------------------------------ */
/* indexing helpers                                                    */
/* ------------------------------------------------------------------ */
static inline long CX(const ob_blk *b, int i, int j, int k) {
  return (long)(i - (1 - GC)) + (long)(j - (1 - GC)) * b->s1 +
         (long)(k - (1 - GC)) * b->s2;
}
static inline long IX(const ob_blk *b, int i, int j, int k) {
  return (long)(i - 1) + (long)b->n[0] * ((long)(j - 1) + (long)b->n[1] * (k - 1));
}
static inline long NX(const ob_blk *b, int i, int j, int k) {
  return (long)i + (long)(b->n[0] + 1) * ((long)j + (long)(b->n[1] + 1) * k);
}
/* Face arrays per Allocate_Block (src/lib/driver/Mod_Allocate_Data.f90:82-84):
 * dir1 f(0:ni,1:nj,1:nk), dir2 f(1:ni,0:nj,1:nk), dir3 f(1:ni,1:nj,0:nk). */
static inline long FX(const ob_blk *b, int d, int i, int j, int k) {
  int lo[3] = {1, 1, 1}, ext[3] = {b->n[0], b->n[1], b->n[2]};
  lo[d] = 0;
  ext[d] += 1;
  return (long)(i - lo[0]) + (long)ext[0] * ((long)(j - lo[1]) + (long)ext[1] * (k - lo[2]));
}
static inline long nfaces(const ob_blk *b, int d) {
  long e[3] = {b->n[0], b->n[1], b->n[2]};
  e[d] += 1;
  return e[0] * e[1] * e[2];
}
static inline double *Mc(const ob_blk *b, int i, int j, int k) {
  return b->M + 9 * CX(b, i, j, k);
}
#define MEL(m, r, c) ((m)[((r)-1) + 3 * ((c)-1)])

static void set_err(ob_ctx *c, int code, int b, int i, int j, int k) {
#ifdef _OPENMP
#pragma omp critical(ob_err)
#endif
  if (!c->err) {
    c->err = code;
    c->eb = b;
    c->ei = i;
    c->ej = j;
    c->ek = k;
  }
}

/* ------------------------------------------------------------------ */
/* physics: src/lib/physics/Lib_Solid.f90                              */
/* ------------------------------------------------------------------ */
static inline int tab_ok(ob_ctx *c, int Tm) { return Tm >= c->Ti1 && Tm + 1 <= c->Ti2; }
/* int(T) of the reference, taken only where T is a representable int (the
 * conversion of NaN or |T| >= 2^31 is undefined in C); such T fail the test. */
static inline int tab_row(ob_ctx *c, double T, int *Tm) {
  if (!(T > -1.0e9 && T < 1.0e9)) return 0;
  *Tm = (int)T;
  return tab_ok(c, *Tm);
}
static inline double tab(const ob_ctx *c, const double *a, double matID, int Ti) {
  int m = (int)matID; /* REAL(8) subscript, truncated */
  return a[(long)(m - 1) * c->nT + (Ti - c->Ti1)];
}
/* co_Kappa/co_Rho/co_Cp/co_H, Lib_Solid.f90:11-76: linear interpolation at
 * Tm=int(T), v = vm + (T-Tm)/(Tp-Tm)*(vp-vm). */
static double co_interp(ob_ctx *c, const double *a, double matID, double T) {
  int Tm = 0;
  if (!tab_row(c, T, &Tm)) {
    set_err(c, OB_ERR_TABLE, 0, 0, 0, 0);
    return NAN;
  }
  const int Tp = Tm + 1;
  double vm = tab(c, a, matID, Tm), vp = tab(c, a, matID, Tp);
  return vm + (T - (double)Tm) / (double)(Tp - Tm) * (vp - vm);
}
#define co_Kappa(c, m, T) co_interp(c, (c)->kap, m, T)
#define co_Rho(c, m, T) co_interp(c, (c)->rho, m, T)
#define co_Cp(c, m, T) co_interp(c, (c)->cp, m, T)
#define co_H(c, m, T) co_interp(c, (c)->h, m, T)
/* co_DrhoDT/co_DcpDT, Lib_Solid.f90:101-132: (vp-vm)/(Tp-Tm). */
static double co_slope(ob_ctx *c, const double *a, double matID, double T) {
  int Tm = 0;
  if (!tab_row(c, T, &Tm)) {
    set_err(c, OB_ERR_TABLE, 0, 0, 0, 0);
    return NAN;
  }
  const int Tp = Tm + 1;
  return (tab(c, a, matID, Tp) - tab(c, a, matID, Tm)) / (double)(Tp - Tm);
}
/* co_T, Lib_Solid.f90:79-98: first subscript i=1..size with h(i) > h, then
 * T = Tm + (h-hm)/(hp-hm)*(Tp-Tm).  The scan is replaced by a binary search
 * for the same first index (h is checked non-decreasing at load). */
static double co_T(ob_ctx *c, double matID, double hv) {
  int m = (int)matID;
  const double *row = c->h + (long)(m - 1) * c->nT;
  /* subscripts scanned by the reference: 1..nT, valid when >= Ti1 */
  int lo = 1, hi = c->nT; /* search first s in [lo,hi] with h(s) > hv */
  if (!(hv == hv)) {
    set_err(c, OB_ERR_TABLE, 0, 0, 0, 0);
    return NAN;
  }
  while (lo < hi) {
    int mid = lo + (hi - lo) / 2;
    if (mid < c->Ti1 || row[mid - c->Ti1] > hv)
      hi = mid;
    else
      lo = mid + 1;
  }
  int s = lo;
  if (!(s >= c->Ti1 + 1 && s <= c->Ti2 && row[s - c->Ti1] > hv)) {
    set_err(c, OB_ERR_TABLE, 0, 0, 0, 0);
    return NAN;
  }
  int Tm = s - 1, Tp = s;
  double hm = row[Tm - c->Ti1], hp = row[Tp - c->Ti1];
  return (double)Tm + (hv - hm) / (hp - hm) * (double)(Tp - Tm);
}

/* ------------------------------------------------------------------ */
/* metrics: src/lib/numerics/space/Lib_Metrics.f90                      */
/* ------------------------------------------------------------------ */
static double tvol(const double *vx, const double *vy, const double *vz, int i1,
                   int i2, int i3, int i4) {
  /* Lib_Metrics.f90:360-374 (1-based vertex ids) */
  i1--; i2--; i3--; i4--;
  return fabs(((vx[i2] - vx[i1]) *
                   ((vy[i3] - vy[i1]) * (vz[i4] - vz[i1]) - (vy[i4] - vy[i1]) * (vz[i3] - vz[i1])) +
               (vy[i2] - vy[i1]) *
                   ((vx[i4] - vx[i1]) * (vz[i3] - vz[i1]) - (vx[i3] - vx[i1]) * (vz[i4] - vz[i1])) +
               (vz[i2] - vz[i1]) *
                   ((vx[i3] - vx[i1]) * (vy[i4] - vy[i1]) - (vx[i4] - vx[i1]) * (vy[i3] - vy[i1]))) /
              6.0);
}

/* Compute_Metric_Tensor, Lib_Metrics.f90:250-376. Returns 0 or error. */
static int metric_tensor(const double *N1, const double *N2, const double *N3,
                         const double *N4, const double *N5, const double *N6,
                         const double *N7, const double *N8, double *M, double *dl,
                         double *vol) {
  double A[3][3], cof[3][3], det;
  for (int x = 0; x < 3; x++) {
    A[0][x] = N5[x] - N1[x] + N6[x] - N2[x] + N7[x] - N3[x] + N8[x] - N4[x];
    A[1][x] = N3[x] - N1[x] + N4[x] - N2[x] + N7[x] - N5[x] + N8[x] - N6[x];
    A[2][x] = N2[x] - N1[x] + N4[x] - N3[x] + N6[x] - N5[x] + N8[x] - N7[x];
  }
  for (int r = 0; r < 3; r++)
    for (int x = 0; x < 3; x++) A[r][x] = 0.25 * A[r][x];
  det = A[0][0] * A[1][1] * A[2][2] - A[0][0] * A[1][2] * A[2][1] -
        A[0][1] * A[1][0] * A[2][2] + A[0][1] * A[1][2] * A[2][0] +
        A[0][2] * A[1][0] * A[2][1] - A[0][2] * A[1][1] * A[2][0];
  if (fabs(det) == 0.0) return OB_ERR_METRIC;
  cof[0][0] = (A[1][1] * A[2][2] - A[1][2] * A[2][1]);
  cof[0][1] = -(A[1][0] * A[2][2] - A[1][2] * A[2][0]);
  cof[0][2] = (A[1][0] * A[2][1] - A[1][1] * A[2][0]);
  cof[1][0] = -(A[0][1] * A[2][2] - A[0][2] * A[2][1]);
  cof[1][1] = (A[0][0] * A[2][2] - A[0][2] * A[2][0]);
  cof[1][2] = -(A[0][0] * A[2][1] - A[0][1] * A[2][0]);
  cof[2][0] = (A[0][1] * A[1][2] - A[0][2] * A[1][1]);
  cof[2][1] = -(A[0][0] * A[1][2] - A[0][2] * A[1][0]);
  cof[2][2] = (A[0][0] * A[1][1] - A[0][1] * A[1][0]);
  for (int r = 0; r < 3; r++)
    for (int x = 0; x < 3; x++) MEL(M, r + 1, x + 1) = cof[r][x] / det;
  for (int h = 0; h < 3; h++)
    dl[h] = sqrt(A[h][0] * A[h][0] + A[h][1] * A[h][1] + A[h][2] * A[h][2]);
  double vx[8], vy[8], vz[8];
  const double *P[8] = {N1, N5, N3, N7, N2, N6, N4, N8};
  for (int q = 0; q < 8; q++) {
    vx[q] = P[q][0];
    vy[q] = P[q][1];
    vz[q] = P[q][2];
  }
  *vol = tvol(vx, vy, vz, 1, 2, 3, 5) + tvol(vx, vy, vz, 2, 4, 3, 8) +
         tvol(vx, vy, vz, 5, 8, 6, 2) + tvol(vx, vy, vz, 5, 7, 8, 3) +
         tvol(vx, vy, vz, 5, 8, 2, 3);
  if (*vol <= 0.0) return OB_ERR_METRIC;
  return 0;
}

static inline const double *ND(const ob_blk *b, int i, int j, int k) {
  return b->node + 3 * NX(b, i, j, k);
}

static void cross(const double *d1, const double *d2, double *d3) {
  d3[0] = (d1[1] * d2[2] - d1[2] * d2[1]);
  d3[1] = (d1[2] * d2[0] - d1[0] * d2[2]);
  d3[2] = (d1[0] * d2[1] - d1[1] * d2[0]);
}
static double sign_of(const ob_blk *b, int dir) {
  /* Compute_Norm_Area sign determination, Lib_Metrics.f90:86-153 */
  double d1[3], d2[3], d3[3], sn[3];
  const double *a, *bb, *cc, *dd;
  if (dir == 0) {
    for (int x = 0; x < 3; x++) {
      d1[x] = ND(b, 1, 1, 1)[x] - ND(b, 1, 0, 0)[x];
      d2[x] = ND(b, 1, 0, 1)[x] - ND(b, 1, 1, 0)[x];
    }
  } else if (dir == 1) {
    for (int x = 0; x < 3; x++) {
      d1[x] = ND(b, 0, 1, 1)[x] - ND(b, 1, 1, 0)[x];
      d2[x] = ND(b, 1, 1, 1)[x] - ND(b, 0, 1, 0)[x];
    }
  } else {
    for (int x = 0; x < 3; x++) {
      d1[x] = ND(b, 0, 1, 1)[x] - ND(b, 1, 0, 1)[x];
      d2[x] = ND(b, 0, 0, 1)[x] - ND(b, 1, 1, 1)[x];
    }
  }
  cross(d1, d2, sn);
  if (dir == 0) {
    a = ND(b, 1, 1, 1); bb = ND(b, 1, 0, 1); cc = ND(b, 1, 0, 0); dd = ND(b, 1, 1, 0);
    for (int x = 0; x < 3; x++) d1[x] = 0.25 * (a[x] + bb[x] + cc[x] + dd[x]);
    a = ND(b, 0, 1, 1); bb = ND(b, 0, 0, 1); cc = ND(b, 0, 0, 0); dd = ND(b, 0, 1, 0);
    for (int x = 0; x < 3; x++) d2[x] = 0.25 * (a[x] + bb[x] + cc[x] + dd[x]);
  } else if (dir == 1) {
    a = ND(b, 1, 1, 1); bb = ND(b, 0, 1, 1); cc = ND(b, 0, 1, 0); dd = ND(b, 1, 1, 0);
    for (int x = 0; x < 3; x++) d1[x] = 0.25 * (a[x] + bb[x] + cc[x] + dd[x]);
    a = ND(b, 1, 0, 1); bb = ND(b, 0, 0, 1); cc = ND(b, 0, 0, 0); dd = ND(b, 1, 0, 0);
    for (int x = 0; x < 3; x++) d2[x] = 0.25 * (a[x] + bb[x] + cc[x] + dd[x]);
  } else {
    a = ND(b, 1, 1, 1); bb = ND(b, 0, 1, 1); cc = ND(b, 0, 0, 1); dd = ND(b, 1, 0, 1);
    for (int x = 0; x < 3; x++) d1[x] = 0.25 * (a[x] + bb[x] + cc[x] + dd[x]);
    a = ND(b, 1, 1, 0); bb = ND(b, 0, 1, 0); cc = ND(b, 0, 0, 0); dd = ND(b, 1, 0, 0);
    for (int x = 0; x < 3; x++) d2[x] = 0.25 * (a[x] + bb[x] + cc[x] + dd[x]);
  }
  for (int x = 0; x < 3; x++) d3[x] = d1[x] - d2[x];
  double scal = d3[0] * sn[0] + d3[1] * sn[1] + d3[2] * sn[2];
  return scal >= 0.0 ? 1.0 : -1.0; /* Fortran sign(1d0, scal) */
}

/* Compute_Norm_Area, Lib_Metrics.f90:69-247 */
static void norm_area(ob_blk *b) {
  double sg[3];
  for (int d = 0; d < 3; d++) sg[d] = sign_of(b, d);
  int im = b->n[0], jm = b->n[1], km = b->n[2];
  for (int d = 0; d < 3; d++) {
    int i0 = d == 0 ? 0 : 1, j0 = d == 1 ? 0 : 1, k0 = d == 2 ? 0 : 1;
    for (int k = k0; k <= km; k++)
      for (int j = j0; j <= jm; j++)
        for (int i = i0; i <= im; i++) {
          double d1[3], d2[3], d3[3];
          const double *p, *q, *r, *s;
          if (d == 0) {
            p = ND(b, i, j, k); q = ND(b, i, j - 1, k - 1); r = ND(b, i, j - 1, k); s = ND(b, i, j, k - 1);
          } else if (d == 1) {
            p = ND(b, i - 1, j, k); q = ND(b, i, j, k - 1); r = ND(b, i, j, k); s = ND(b, i - 1, j, k - 1);
          } else {
            p = ND(b, i - 1, j, k); q = ND(b, i, j - 1, k); r = ND(b, i - 1, j - 1, k); s = ND(b, i, j, k);
          }
          for (int x = 0; x < 3; x++) {
            d1[x] = p[x] - q[x];
            d2[x] = r[x] - s[x];
          }
          d3[0] = (d1[1] * d2[2] - d1[2] * d2[1]) * .5;
          d3[1] = (d1[2] * d2[0] - d1[0] * d2[2]) * .5;
          d3[2] = (d1[0] * d2[1] - d1[1] * d2[0]) * .5;
          double A = sqrt(d3[0] * d3[0] + d3[1] * d3[1] + d3[2] * d3[2]);
          double nx = d3[0] / A * sg[d], ny = d3[1] / A * sg[d], nz = d3[2] / A * sg[d];
          if (A == 0.0) nx = ny = nz = 0.0;
          long f = FX(b, d, i, j, k);
          b->fA[d][f] = A;
          b->fN[d][3 * f + 0] = nx;
          b->fN[d][3 * f + 1] = ny;
          b->fN[d][3 * f + 2] = nz;
        }
  }
}

/* Setup_Metrics interior part, src/lib/numerics/space/Mod_Metrics.f90:23-45 */
static int setup_block_metrics(ob_ctx *c, ob_blk *b) {
  for (int k = 1; k <= b->n[2]; k++)
    for (int j = 1; j <= b->n[1]; j++)
      for (int i = 1; i <= b->n[0]; i++) {
        long p = CX(b, i, j, k);
        int e = metric_tensor(ND(b, i - 1, j - 1, k - 1), ND(b, i - 1, j - 1, k),
                              ND(b, i - 1, j, k - 1), ND(b, i - 1, j, k),
                              ND(b, i, j - 1, k - 1), ND(b, i, j - 1, k),
                              ND(b, i, j, k - 1), ND(b, i, j, k), b->M + 9 * p,
                              b->dl + 3 * p, b->vol + p);
        if (e) {
          set_err(c, OB_ERR_METRIC, 0, i, j, k);
          return e;
        }
      }
  norm_area(b);
  return 0;
}

/* BC_Connect_Metrics, Lib_Metrics.f90:704-936: ghost metric Mg(1:2) of a
 * 101 entry from the local face nodes and the donor block's nodes. */
static int connect_metrics(ob_ctx *c, ob_bc *e, ob_blk *bm, ob_blk *bs) {
  static const int gs[6][3] = {{0, 0, 0}, {-1, 0, 0}, {0, 0, 0}, {0, -1, 0}, {0, 0, 0}, {0, 0, -1}};
  static const int gs2[6][3] = {{1, 0, 0}, {-2, 0, 0}, {0, 1, 0}, {0, -2, 0}, {0, 0, 1}, {0, 0, -2}};
  double N[8][GC + 1][3];
  int Im = e->i, Jm = e->j, Km = e->k, Fm = e->f, Fs = e->fs;
  int d11 = e->d11, d12 = e->d12, d21 = e->d21, d22 = e->d22;
  memset(N, 0, sizeof(N));
#define SETN(q, ii, jj, kk) memcpy(N[(q)-1][0], ND(bm, ii, jj, kk), 3 * sizeof(double))
  switch (Fm) {
    case 1: SETN(1, Im - 1, Jm - 1, Km - 1); SETN(2, Im - 1, Jm - 1, Km); SETN(3, Im - 1, Jm, Km - 1); SETN(4, Im - 1, Jm, Km); break;
    case 2: SETN(5, Im, Jm - 1, Km - 1); SETN(6, Im, Jm - 1, Km); SETN(7, Im, Jm, Km - 1); SETN(8, Im, Jm, Km); break;
    case 3: SETN(1, Im - 1, Jm - 1, Km - 1); SETN(2, Im - 1, Jm - 1, Km); SETN(5, Im, Jm - 1, Km - 1); SETN(6, Im, Jm - 1, Km); break;
    case 4: SETN(3, Im - 1, Jm, Km - 1); SETN(4, Im - 1, Jm, Km); SETN(7, Im, Jm, Km - 1); SETN(8, Im, Jm, Km); break;
    case 5: SETN(1, Im - 1, Jm - 1, Km - 1); SETN(3, Im - 1, Jm, Km - 1); SETN(5, Im, Jm - 1, Km - 1); SETN(7, Im, Jm, Km - 1); break;
    case 6: SETN(2, Im - 1, Jm - 1, Km); SETN(4, Im - 1, Jm, Km); SETN(6, Im, Jm - 1, Km); SETN(8, Im, Jm, Km); break;
  }
#undef SETN
  for (int g = 1; g <= GC; g++) {
    int Is1 = e->is + (GC - g) * gs[Fs - 1][0] + (g - 1) * gs2[Fs - 1][0];
    int Js1 = e->js + (GC - g) * gs[Fs - 1][1] + (g - 1) * gs2[Fs - 1][1];
    int Ks1 = e->ks + (GC - g) * gs[Fs - 1][2] + (g - 1) * gs2[Fs - 1][2];
    int II, JJ, KK;
    if (Fs <= 2) { II = Is1; JJ = (2 * Js1 + d11 + d21) / 2; KK = (2 * Ks1 + d12 + d22) / 2; }
    else if (Fs <= 4) { II = (2 * Is1 + d11 + d21) / 2; JJ = Js1; KK = (2 * Ks1 + d12 + d22) / 2; }
    else { II = (2 * Is1 + d11 + d21) / 2; JJ = (2 * Js1 + d12 + d22) / 2; KK = Ks1; }
    int III, JJJ, KKK;
    static const int q1[6] = {4, 8, 6, 8, 7, 8}, q2[6] = {2, 6, 2, 4, 3, 4},
                     q3[6] = {3, 7, 5, 7, 5, 6}, q4[6] = {1, 5, 1, 3, 1, 2};
    if (II < 0 || II > bs->n[0] || JJ < 0 || JJ > bs->n[1] || KK < 0 || KK > bs->n[2])
      return OB_ERR_ARG;
    memcpy(N[q1[Fm - 1] - 1][g], ND(bs, II, JJ, KK), 3 * sizeof(double));
    /* node (i1-1, i2) */
    if (Fs <= 2) { III = II; JJJ = JJ - d11; KKK = KK - d12; }
    else if (Fs <= 4) { III = II - d11; JJJ = JJ; KKK = KK - d12; }
    else { III = II - d11; JJJ = JJ - d12; KKK = KK; }
    if (III < 0 || III > bs->n[0] || JJJ < 0 || JJJ > bs->n[1] || KKK < 0 || KKK > bs->n[2]) return OB_ERR_ARG;
    memcpy(N[q2[Fm - 1] - 1][g], ND(bs, III, JJJ, KKK), 3 * sizeof(double));
    /* node (i1, i2-1) */
    if (Fs <= 2) { III = II; JJJ = JJ - d21; KKK = KK - d22; }
    else if (Fs <= 4) { III = II - d21; JJJ = JJ; KKK = KK - d22; }
    else { III = II - d21; JJJ = JJ - d22; KKK = KK; }
    if (III < 0 || III > bs->n[0] || JJJ < 0 || JJJ > bs->n[1] || KKK < 0 || KKK > bs->n[2]) return OB_ERR_ARG;
    memcpy(N[q3[Fm - 1] - 1][g], ND(bs, III, JJJ, KKK), 3 * sizeof(double));
    /* node (i1-1, i2-1) */
    if (Fs <= 2) { III = II; JJJ = JJ - d11 - d21; KKK = KK - d12 - d22; }
    else if (Fs <= 4) { III = II - d11 - d21; JJJ = JJ; KKK = KK - d12 - d22; }
    else { III = II - d11 - d21; JJJ = JJ - d12 - d22; KKK = KK; }
    if (III < 0 || III > bs->n[0] || JJJ < 0 || JJJ > bs->n[1] || KKK < 0 || KKK > bs->n[2]) return OB_ERR_ARG;
    memcpy(N[q4[Fm - 1] - 1][g], ND(bs, III, JJJ, KKK), 3 * sizeof(double));
    double dl[3], vol;
    double *Mg = e->Mg[g - 1];
    int err = 0;
#define P(q, gg) N[(q)-1][gg]
    switch (Fm) {
      case 1: err = metric_tensor(P(1, g), P(2, g), P(3, g), P(4, g), P(1, g - 1), P(2, g - 1), P(3, g - 1), P(4, g - 1), Mg, dl, &vol); break;
      case 2: err = metric_tensor(P(5, g - 1), P(6, g - 1), P(7, g - 1), P(8, g - 1), P(5, g), P(6, g), P(7, g), P(8, g), Mg, dl, &vol); break;
      case 3: err = metric_tensor(P(1, g), P(2, g), P(1, g - 1), P(2, g - 1), P(5, g), P(6, g), P(5, g - 1), P(6, g - 1), Mg, dl, &vol); break;
      case 4: err = metric_tensor(P(3, g - 1), P(4, g - 1), P(3, g), P(4, g), P(7, g - 1), P(8, g - 1), P(7, g), P(8, g), Mg, dl, &vol); break;
      case 5: err = metric_tensor(P(1, g), P(1, g - 1), P(3, g), P(3, g - 1), P(5, g), P(5, g - 1), P(7, g), P(7, g - 1), Mg, dl, &vol); break;
      case 6: err = metric_tensor(P(2, g - 1), P(2, g), P(4, g - 1), P(4, g), P(6, g - 1), P(6, g), P(8, g - 1), P(8, g), Mg, dl, &vol); break;
    }
#undef P
    if (err) {
      set_err(c, OB_ERR_METRIC, e->b, e->i, e->j, e->k);
      return err;
    }
  }
  return 0;
}

/* BC_Extrapolate_Metrics, Lib_Metrics.f90:379-600, first ghost layer: the
 * ghost cell's nodes extrapolated along the face normal from the block's
 * node layers (cubic, 3a - 3b + c; linear, 2a - b, across j faces of blocks
 * with 2 or fewer j cells and across k faces in 2-D), its metric tensor
 * Mg(1).  Only chimera entries read it (BC_Connection's M_con). */
static void xnode(double *out, const double *a, const double *b, const double *cc, int cubic) {
  for (int x = 0; x < 3; x++) out[x] = cubic ? 3.0 * a[x] - 3.0 * b[x] + cc[x] : 2.0 * a[x] - b[x];
}
static int extrap_metric(ob_ctx *c, ob_bc *e, ob_blk *bm) {
  int Im = e->i, Jm = e->j, Km = e->k, Fm = e->f;
  /* per face: the 4 node columns (i, j, k offsets of layer 0), the step
   * into the block (-1: the ghost lies at +1), whether the ghost is on the
   * min side, cubic or linear */
  int cubic = 1;
  if ((Fm == 3 || Fm == 4) && bm->n[1] <= 2) cubic = 0;
  if ((Fm == 5 || Fm == 6) && c->ndir != 3) cubic = 0;
  int base[4][3], di = 0, dj = 0, dk = 0;
  switch (Fm) {
    case 1: { int q[4][3] = {{Im - 1, Jm - 1, Km - 1}, {Im - 1, Jm - 1, Km}, {Im - 1, Jm, Km - 1}, {Im - 1, Jm, Km}}; memcpy(base, q, sizeof(base)); di = 1; } break;
    case 2: { int q[4][3] = {{Im, Jm - 1, Km - 1}, {Im, Jm - 1, Km}, {Im, Jm, Km - 1}, {Im, Jm, Km}}; memcpy(base, q, sizeof(base)); di = -1; } break;
    case 3: { int q[4][3] = {{Im - 1, Jm - 1, Km - 1}, {Im - 1, Jm - 1, Km}, {Im, Jm - 1, Km - 1}, {Im, Jm - 1, Km}}; memcpy(base, q, sizeof(base)); dj = 1; } break;
    case 4: { int q[4][3] = {{Im - 1, Jm, Km - 1}, {Im - 1, Jm, Km}, {Im, Jm, Km - 1}, {Im, Jm, Km}}; memcpy(base, q, sizeof(base)); dj = -1; } break;
    case 5: { int q[4][3] = {{Im - 1, Jm - 1, Km - 1}, {Im - 1, Jm, Km - 1}, {Im, Jm - 1, Km - 1}, {Im, Jm, Km - 1}}; memcpy(base, q, sizeof(base)); dk = 1; } break;
    default: { int q[4][3] = {{Im - 1, Jm - 1, Km}, {Im - 1, Jm, Km}, {Im, Jm - 1, Km}, {Im, Jm, Km}}; memcpy(base, q, sizeof(base)); dk = -1; } break;
  }
  double L0[4][3], G[4][3];
  for (int q = 0; q < 4; q++) {
    int i0 = base[q][0], j0 = base[q][1], k0 = base[q][2];
    for (int s2 = 0; s2 <= 1 + cubic; s2++) {
      int ii = i0 + di * s2, jj = j0 + dj * s2, kk = k0 + dk * s2;
      if (ii < 0 || ii > bm->n[0] || jj < 0 || jj > bm->n[1] || kk < 0 || kk > bm->n[2]) return OB_ERR_ARG;
    }
    const double *a = ND(bm, i0, j0, k0), *b = ND(bm, i0 + di, j0 + dj, k0 + dk),
                 *cc = cubic ? ND(bm, i0 + 2 * di, j0 + 2 * dj, k0 + 2 * dk) : a;
    memcpy(L0[q], a, 3 * sizeof(double));
    xnode(G[q], a, b, cc, cubic);
  }
  double dl[3], vol;
  double *Mg = e->Mg[0];
  int err = 0;
  switch (Fm) {
    case 1: err = metric_tensor(G[0], G[1], G[2], G[3], L0[0], L0[1], L0[2], L0[3], Mg, dl, &vol); break;
    case 2: err = metric_tensor(L0[0], L0[1], L0[2], L0[3], G[0], G[1], G[2], G[3], Mg, dl, &vol); break;
    case 3: err = metric_tensor(G[0], G[1], L0[0], L0[1], G[2], G[3], L0[2], L0[3], Mg, dl, &vol); break;
    case 4: err = metric_tensor(L0[0], L0[1], G[0], G[1], L0[2], L0[3], G[2], G[3], Mg, dl, &vol); break;
    case 5: err = metric_tensor(G[0], L0[0], G[1], L0[1], G[2], L0[2], G[3], L0[3], Mg, dl, &vol); break;
    default: err = metric_tensor(L0[0], G[0], L0[1], G[1], L0[2], G[2], L0[3], G[3], Mg, dl, &vol); break;
  }
  if (err) set_err(c, OB_ERR_METRIC, e->b, e->i, e->j, e->k);
  return err;
}

/* ------------------------------------------------------------------ */
/* ghost fill: src/lib/numerics/space/Lib_Ghost.f90                     */
/* ------------------------------------------------------------------ */
static void ghost_fill(ob_ctx *c, ob_dom *d) {
  /* local BC loop, Lib_Ghost.f90:42-68 */
  for (int n = 0; n < d->nbound; n++) {
    ob_bc *e = &d->bc[n];
    ob_blk *bm = &d->blk[e->b - 1];
    int Fm = e->f;
    if (e->type == 101) {
      /* Ghost_Connection, Lib_Ghost.f90:132-156 */
      ob_blk *bs = &d->blk[e->bs - 1];
      for (int g = 1; g <= GC; g++) {
        int Ig = e->i - GUIDE[Fm - 1][0] * g, Jg = e->j - GUIDE[Fm - 1][1] * g,
            Kg = e->k - GUIDE[Fm - 1][2] * g;
        int Is = e->is + GUIDE[e->fs - 1][0] * (g - 1), Js = e->js + GUIDE[e->fs - 1][1] * (g - 1),
            Ks = e->ks + GUIDE[e->fs - 1][2] * (g - 1);
        bm->T[CX(bm, Ig, Jg, Kg)] = bs->T[CX(bs, Is, Js, Ks)];
        bm->matID[CX(bm, Ig, Jg, Kg)] = bs->matID[CX(bs, Is, Js, Ks)];
        e->Tg[g - 1] = bm->T[CX(bm, Ig, Jg, Kg)];
        if (g == 1) e->matIDg = bm->matID[CX(bm, Ig, Jg, Kg)];
      }
    } else if (e->type == 102) {
      /* Ghost_Chimera, Lib_Ghost.f90:192-245: the first ghost layer from the
       * volume-fraction-weighted enthalpy of its donors, inverted with the
       * material of the last donor.  The second layer (donors nch1+1..nch)
       * is not formed: no term of the step reads it, and the reference
       * inverts it with a ghost material that nothing sets. */
      int Ig = e->i - GUIDE[Fm - 1][0], Jg = e->j - GUIDE[Fm - 1][1], Kg = e->k - GUIDE[Fm - 1][2];
      long pg = CX(bm, Ig, Jg, Kg);
      double consg = 0.0;
      for (int q = 0; q < e->nch1; q++) {
        const int *dn = e->cdon + 4 * q;
        ob_blk *bs = &d->blk[dn[0] - 1];
        long ps = CX(bs, dn[1], dn[2], dn[3]);
        double consi = co_H(c, bs->matID[ps], bs->T[ps]);
        consg = consg + consi * e->cvf[q];
        bm->matID[pg] = bs->matID[ps];
      }
      bm->T[pg] = co_T(c, bm->matID[pg], consg);
      e->Tg[0] = bm->T[pg];
      e->matIDg = bm->matID[pg];
    } else {
      /* Ghost_Symmetry (also the default for walls/null), Lib_Ghost.f90:159-189 */
      for (int g = 1; g <= GC; g++) {
        int gi = GUIDE[Fm - 1][0], gj = GUIDE[Fm - 1][1], gk = GUIDE[Fm - 1][2];
        long dst = CX(bm, e->i - gi * g, e->j - gj * g, e->k - gk * g);
        long src = CX(bm, e->i + gi * (g - 1), e->j + gj * (g - 1), e->k + gk * (g - 1));
        bm->T[dst] = bm->T[src];
        bm->matID[dst] = bm->matID[src];
      }
    }
  }
  /* Tg(3:6) of connections, Lib_Ghost.f90:72-86 with Fill_BC_Ghost_Connection :296-351 */
  for (int n = 0; n < d->nbound; n++) {
    ob_bc *e = &d->bc[n];
    if (e->type != 101) continue;
    ob_blk *bs = &d->blk[e->bs - 1];
    int Is = e->is, Js = e->js, Ks = e->ks;
    int p[4][3];
    if (e->fs <= 2) {
      int q[4][3] = {{Is, Js - e->d11, Ks - e->d12}, {Is, Js + e->d11, Ks + e->d12},
                     {Is, Js - e->d21, Ks - e->d22}, {Is, Js + e->d21, Ks + e->d22}};
      memcpy(p, q, sizeof(p));
    } else if (e->fs <= 4) {
      int q[4][3] = {{Is - e->d11, Js, Ks - e->d12}, {Is + e->d11, Js, Ks + e->d12},
                     {Is - e->d21, Js, Ks - e->d22}, {Is + e->d21, Js, Ks + e->d22}};
      memcpy(p, q, sizeof(p));
    } else {
      int q[4][3] = {{Is - e->d11, Js - e->d12, Ks}, {Is + e->d11, Js + e->d12, Ks},
                     {Is - e->d21, Js - e->d22, Ks}, {Is + e->d21, Js + e->d22, Ks}};
      memcpy(p, q, sizeof(p));
    }
    for (int q = 0; q < 4; q++) e->Tg[2 + q] = bs->T[CX(bs, p[q][0], p[q][1], p[q][2])];
  }
  /* Tg(3:6) of chimera entries, Lib_Ghost.f90:88-101 with
   * Fill_BC_Ghost_Chimera :354-433: the ghost's tangential neighbours at the
   * block's edges take its value, then Tg(3:6) are the boundary cell's own
   * tangential neighbours */
  for (int n = 0; n < d->nbound; n++) {
    ob_bc *e = &d->bc[n];
    if (e->type != 102) continue;
    ob_blk *bm = &d->blk[e->b - 1];
    int Im = e->i, Jm = e->j, Km = e->k, Fm = e->f;
    int Ig = Im - GUIDE[Fm - 1][0], Jg = Jm - GUIDE[Fm - 1][1], Kg = Km - GUIDE[Fm - 1][2];
    int *n3 = bm->n;
    int q[4][3];
    if (Fm <= 2) {
      if (Jm == 1) bm->T[CX(bm, Ig, Jm - 1, Km)] = bm->T[CX(bm, Ig, Jm, Km)];
      if (Jm == n3[1]) bm->T[CX(bm, Ig, Jm + 1, Km)] = bm->T[CX(bm, Ig, Jm, Km)];
      if (Km == 1) bm->T[CX(bm, Ig, Jm, Km - 1)] = bm->T[CX(bm, Ig, Jm, Km)];
      if (Km == n3[2]) bm->T[CX(bm, Ig, Jm, Km + 1)] = bm->T[CX(bm, Ig, Jm, Km)];
      int r[4][3] = {{Im, Jm - 1, Km}, {Im, Jm + 1, Km}, {Im, Jm, Km - 1}, {Im, Jm, Km + 1}};
      memcpy(q, r, sizeof(q));
    } else if (Fm <= 4) {
      if (Im == 1) bm->T[CX(bm, Im - 1, Jg, Km)] = bm->T[CX(bm, Im, Jg, Km)];
      if (Im == n3[0]) bm->T[CX(bm, Im + 1, Jg, Km)] = bm->T[CX(bm, Im, Jg, Km)];
      if (Km == 1) bm->T[CX(bm, Im, Jg, Km - 1)] = bm->T[CX(bm, Im, Jg, Km)];
      if (Km == n3[2]) bm->T[CX(bm, Im, Jg, Km + 1)] = bm->T[CX(bm, Im, Jg, Km)];
      int r[4][3] = {{Im - 1, Jm, Km}, {Im + 1, Jm, Km}, {Im, Jm, Km - 1}, {Im, Jm, Km + 1}};
      memcpy(q, r, sizeof(q));
    } else {
      if (Im == 1) bm->T[CX(bm, Im - 1, Jm, Kg)] = bm->T[CX(bm, Im, Jm, Kg)];
      if (Im == n3[0]) bm->T[CX(bm, Im + 1, Jm, Kg)] = bm->T[CX(bm, Im, Jm, Kg)];
      if (Jm == 1) bm->T[CX(bm, Im, Jm - 1, Kg)] = bm->T[CX(bm, Im, Jm, Kg)];
      if (Jm == n3[1]) bm->T[CX(bm, Im, Jm + 1, Kg)] = bm->T[CX(bm, Im, Jm, Kg)];
      int r[4][3] = {{Im - 1, Jm, Km}, {Im + 1, Jm, Km}, {Im, Jm - 1, Km}, {Im, Jm + 1, Km}};
      memcpy(q, r, sizeof(q));
    }
    for (int t = 0; t < 4; t++) e->Tg[2 + t] = bm->T[CX(bm, q[t][0], q[t][1], q[t][2])];
  }
}

/* Fill_matIDg, Lib_Ghost.f90:436-470 */
static void fill_matIDg(ob_dom *d) {
  for (int n = 0; n < d->nbound; n++) {
    ob_bc *e = &d->bc[n];
    if (e->type == 101 || e->type == 102) {
      ob_blk *bm = &d->blk[e->b - 1];
      e->matIDg = bm->matID[CX(bm, e->i - GUIDE[e->f - 1][0], e->j - GUIDE[e->f - 1][1],
                               e->k - GUIDE[e->f - 1][2])];
    }
  }
}

/* ------------------------------------------------------------------ */
/* fluxes: src/lib/numerics/fluxes/Lib_Diffusive.f90, Mod_Fluxes.f90    */
/* ------------------------------------------------------------------ */
static inline void matvec_GM(const double G[3], const double *M, double out[3]) {
  /* matmul(Gradient, M): out(x) = (G1*M(1,x) + G2*M(2,x)) + G3*M(3,x) */
  for (int x = 1; x <= 3; x++)
    out[x - 1] = G[0] * MEL(M, 1, x) + G[1] * MEL(M, 2, x) + G[2] * MEL(M, 3, x);
}
static inline double dot3(const double *a, const double *b) {
  return a[0] * b[0] + a[1] * b[1] + a[2] * b[2];
}

/* Diffusive_Flux, Lib_Diffusive.f90:10-45 (aa,bb,cc 1-based) */
static void diffusive_flux(ob_ctx *c, double matID, const double *normal, double area,
                           double T1, double T2, double T3, double T4, double T5,
                           double T6, double T7, double T8, double T9, double T10,
                           const double *M1, const double *M2, double *Res1,
                           double *Res2, int aa, int bb, int cc) {
  double G[3], M[9], gr[3];
  G[aa - 1] = T2 - T1;
  G[bb - 1] = (T4 - T3 + T6 - T5) * 0.25;
  G[cc - 1] = (T8 - T7 + T10 - T9) * 0.25;
  if (c->ndir == 2) G[cc - 1] = 0.0;
  for (int q = 0; q < 9; q++) M[q] = 0.5 * (M1[q] + M2[q]);
  matvec_GM(G, M, gr);
  double T = 0.5 * (T1 + T2);
  double kappa = co_Kappa(c, matID, T);
  double Flux = area * kappa * dot3(gr, normal);
  *Res1 = *Res1 - Flux;
  *Res2 = *Res2 + Flux;
}

/* Fluxes_blk, Mod_Fluxes.f90:31-126 */
static void fluxes_blk(ob_ctx *c, ob_blk *b) {
  int n1 = b->n[0], n2 = b->n[1], n3 = b->n[2];
  double *T = b->T, *R = b->R;
#ifdef _OPENMP
#pragma omp parallel for collapse(2)
#endif
  for (int k = 1; k <= n3; k++)
    for (int j = 1; j <= n2; j++)
      for (int i = 1; i <= n1; i++) R[CX(b, i, j, k)] = 0.0;
#define TT(i, j, k) T[CX(b, i, j, k)]
  /* (OpenMP build: rows, columns and pillars of one sweep touch disjoint
   * cells, and each cell keeps its serial accumulation order) */
#ifdef _OPENMP
#pragma omp parallel for collapse(2)
#endif
  for (int k = 1; k <= n3; k++)
    for (int j = 1; j <= n2; j++)
      for (int i = 1; i <= n1 - 1; i++) {
        long f = FX(b, 0, i, j, k);
        diffusive_flux(c, b->matID[CX(b, i, j, k)], b->fN[0] + 3 * f, b->fA[0][f], TT(i, j, k),
                       TT(i + 1, j, k), TT(i, j - 1, k), TT(i, j + 1, k), TT(i + 1, j - 1, k),
                       TT(i + 1, j + 1, k), TT(i, j, k - 1), TT(i, j, k + 1), TT(i + 1, j, k - 1),
                       TT(i + 1, j, k + 1), Mc(b, i, j, k), Mc(b, i + 1, j, k),
                       &R[CX(b, i, j, k)], &R[CX(b, i + 1, j, k)], 1, 2, 3);
      }
#ifdef _OPENMP
#pragma omp parallel for collapse(2)
#endif
  for (int k = 1; k <= n3; k++)
    for (int i = 1; i <= n1; i++)
      for (int j = 1; j <= n2 - 1; j++) {
        long f = FX(b, 1, i, j, k);
        diffusive_flux(c, b->matID[CX(b, i, j, k)], b->fN[1] + 3 * f, b->fA[1][f], TT(i, j, k),
                       TT(i, j + 1, k), TT(i - 1, j, k), TT(i + 1, j, k), TT(i - 1, j + 1, k),
                       TT(i + 1, j + 1, k), TT(i, j, k - 1), TT(i, j, k + 1), TT(i, j + 1, k - 1),
                       TT(i, j + 1, k + 1), Mc(b, i, j, k), Mc(b, i, j + 1, k),
                       &R[CX(b, i, j, k)], &R[CX(b, i, j + 1, k)], 2, 1, 3);
      }
#ifdef _OPENMP
#pragma omp parallel for collapse(2)
#endif
  for (int j = 1; j <= n2; j++)
    for (int i = 1; i <= n1; i++)
      for (int k = 1; k <= n3 - 1; k++) {
        long f = FX(b, 2, i, j, k);
        diffusive_flux(c, b->matID[CX(b, i, j, k)], b->fN[2] + 3 * f, b->fA[2][f], TT(i, j, k),
                       TT(i, j, k + 1), TT(i - 1, j, k), TT(i + 1, j, k), TT(i - 1, j, k + 1),
                       TT(i + 1, j, k + 1), TT(i, j - 1, k), TT(i, j + 1, k), TT(i, j - 1, k + 1),
                       TT(i, j + 1, k + 1), Mc(b, i, j, k), Mc(b, i, j, k + 1),
                       &R[CX(b, i, j, k)], &R[CX(b, i, j, k + 1)], 3, 1, 2);
      }
#undef TT
}

/* ------------------------------------------------------------------ */
/* boundary fluxes: src/lib/numerics/fluxes/bc/ files                  */
/* ------------------------------------------------------------------ */
/* Face_Index, Lib_BC_Fluxes.f90:27-45 */
static void face_index(int face, int ci, int cj, int ck, int *dir, int *fi, int *fj, int *fk) {
  *fi = ci; *fj = cj; *fk = ck;
  if (face <= 2) { *dir = 1; *fi = ci - face % 2; }
  else if (face <= 4) { *dir = 2; *fj = cj - face % 2; }
  else { *dir = 3; *fk = ck - face % 2; }
}

static double series_update(const ob_series *s, double time) {
  /* time_series_type%update, src/lib/base/Series_Data_m.f90:55-86 */
  double t_ = time, out = 0.0;
  if (s->periodic) {
    double period = s->t[s->n - 1] - s->t[0];
    if (period > 0.0) {
      double a = t_ - s->t[0];
      double m = a - floor(a / period) * period; /* Fortran modulo */
      t_ = s->t[0] + m;
    }
  }
  if (t_ == s->t[0]) return s->v[0];
  for (int i = 1; i < s->n; i++)
    if (t_ > s->t[i - 1] && t_ <= s->t[i])
      return (s->v[i] - s->v[i - 1]) / (s->t[i] - s->t[i - 1]) * (t_ - s->t[i - 1]) + s->v[i - 1];
  if (t_ > s->t[s->n - 1]) out = s->v[s->n - 1];
  return out;
}

static void bc_fluxes(ob_ctx *c, ob_dom *d) {
  /* BC_Fluxes, Mod_BC_Fluxes.f90:10-86; entries are already grouped by
   * (block, face) so a plain pass over the list keeps the reference order. */
  const double HUGE4 = 3.4028234663852886e38; /* huge(0.0) */
  for (int n = 0; n < d->nbound; n++) {
    ob_bc *e = &d->bc[n];
    ob_blk *B = &d->blk[e->b - 1];
    int Im = e->i, Jm = e->j, Km = e->k, Fm = e->f;
    int dir, fi, fj, fk;
    face_index(Fm, Im, Jm, Km, &dir, &fi, &fj, &fk);
    long f = FX(B, dir - 1, fi, fj, fk);
    const double *Nrm = B->fN[dir - 1] + 3 * f;
    double Area = B->fA[dir - 1][f];
    const double *M = Mc(B, Im, Jm, Km);
    long p = CX(B, Im, Jm, Km);
    double *R = B->R;
    int modfm = Fm % 2, modfm2 = 1 - 2 * modfm, modfm3 = 2 * (2 * modfm - 1);
    double *TT = B->T;
    double Tc = TT[p];
    double mat = B->matID[p];
    switch (e->type) {
      case 101:
      case 102: {
        /* BC_Connection, Lib_BC_Fluxes_Connection.f90:12-113 (chimera:
         * Mod_BC_Fluxes.f90:45, with its ghost, Tg and extrapolated Mg) */
        const double *Mcn = e->Mg[0];
        double Tip = TT[CX(B, Im + 1, Jm, Km)], Tim = TT[CX(B, Im - 1, Jm, Km)];
        double Tjp = TT[CX(B, Im, Jm + 1, Km)], Tjm = TT[CX(B, Im, Jm - 1, Km)];
        double Tkp, Tkm;
        if (c->ndir == 2) { Tkp = HUGE4; Tkm = HUGE4; }
        else { Tkp = TT[CX(B, Im, Jm, Km + 1)]; Tkm = TT[CX(B, Im, Jm, Km - 1)]; }
        double Tcon = e->Tg[0];
        double kl = co_Kappa(c, mat, Tc), kc = co_Kappa(c, e->matIDg, Tcon);
        double ml[3], mc[3];
        for (int r = 1; r <= 3; r++) {
          ml[r - 1] = MEL(M, r, 1) * Nrm[0] + MEL(M, r, 2) * Nrm[1] + MEL(M, r, 3) * Nrm[2];
          mc[r - 1] = MEL(Mcn, r, 1) * Nrm[0] + MEL(Mcn, r, 2) * Nrm[1] + MEL(Mcn, r, 3) * Nrm[2];
        }
        double Dl, Nl, Dc, Nc;
        const double *Tg = e->Tg;
        if (Fm <= 2) {
          Dl = 2.0 * ml[0] * modfm2;
          Nl = 2.0 * Tc * ml[0] * modfm2 - 0.5 * (Tjp - Tjm) * ml[1] - 0.5 * (Tkp - Tkm) * ml[2];
          Dc = 2.0 * mc[0] * modfm2;
          Nc = 2.0 * Tcon * mc[0] * modfm2 + 0.5 * (Tg[3] - Tg[2]) * mc[1] + 0.5 * (Tg[5] - Tg[4]) * mc[2];
        } else if (Fm <= 4) {
          Dl = 2.0 * ml[1] * modfm2;
          Nl = -0.5 * (Tip - Tim) * ml[0] + 2.0 * Tc * ml[1] * modfm2 - 0.5 * (Tkp - Tkm) * ml[2];
          Dc = 2.0 * mc[1] * modfm2;
          Nc = 0.5 * (Tg[3] - Tg[2]) * mc[0] + 2.0 * Tcon * mc[1] * modfm2 + 0.5 * (Tg[5] - Tg[4]) * mc[2];
        } else {
          Dl = 2.0 * ml[2] * modfm2;
          Nl = -0.5 * (Tip - Tim) * ml[0] - 0.5 * (Tjp - Tjm) * ml[1] + 2.0 * Tc * ml[2] * modfm2;
          Dc = 2.0 * mc[2] * modfm2;
          Nc = 0.5 * (Tg[3] - Tg[2]) * mc[0] + 0.5 * (Tg[5] - Tg[4]) * mc[1] + 2.0 * Tcon * mc[2] * modfm2;
        }
        double Tint = (kl * Nl + kc * Nc) / (kl * Dl + kc * Dc);
        double G[3], gr[3];
        if (Fm <= 2) { G[0] = 2.0 * (Tint - Tc) * modfm2; G[1] = (Tjp - Tjm) * 0.5; G[2] = (Tkp - Tkm) * 0.5; }
        else if (Fm <= 4) { G[0] = (Tip - Tim) * 0.5; G[1] = 2.0 * (Tint - Tc) * modfm2; G[2] = (Tkp - Tkm) * 0.5; }
        else { G[0] = (Tip - Tim) * 0.5; G[1] = (Tjp - Tjm) * 0.5; G[2] = 2.0 * (Tint - Tc) * modfm2; }
        if (c->ndir == 2) G[2] = 0.0;
        matvec_GM(G, M, gr);
        double Flux = Area * kl * dot3(gr, Nrm);
        R[p] = R[p] - modfm2 * Flux;
      } break;
      case 300: {
        /* BC_Symmetry, Lib_BC_Fluxes_Symmetry.f90:12-67 */
        double Tip = TT[CX(B, Im + 1, Jm, Km)], Tim = TT[CX(B, Im - 1, Jm, Km)];
        double Tjp = TT[CX(B, Im, Jm + 1, Km)], Tjm = TT[CX(B, Im, Jm - 1, Km)];
        double Tkp = TT[CX(B, Im, Jm, Km + 1)], Tkm = TT[CX(B, Im, Jm, Km - 1)];
        double G[3], gr[3];
        if (Fm <= 2) { G[0] = 0.0; G[1] = 0.5 * (Tjp - Tjm) * modfm2; G[2] = 0.5 * (Tkp - Tkm) * modfm2; }
        else if (Fm <= 4) { G[0] = 0.5 * (Tip - Tim) * modfm2; G[1] = 0.0; G[2] = 0.5 * (Tkp - Tkm) * modfm2; }
        else { G[0] = 0.5 * (Tip - Tim) * modfm2; G[1] = 0.5 * (Tjp - Tjm) * modfm2; G[2] = 0.0; }
        if (c->ndir == 2) G[2] = 0.0;
        matvec_GM(G, M, gr);
        double kappa = co_Kappa(c, mat, Tc);
        double Flux = Area * kappa * dot3(gr, Nrm);
        R[p] = R[p] - modfm2 * Flux;
      } break;
      case 301: {
        /* BC_Wall_Heat, Lib_BC_Fluxes_Wall_Heat.f90:12-54 */
        if (e->series >= 0) e->qw = series_update(&c->series[e->series], d->time);
        double Flux = Area * e->qw;
        R[p] = R[p] - modfm2 * Flux;
      } break;
      case 302: {
        /* BC_Wall_Temperature, Lib_BC_Fluxes_Wall_Temperature.f90:12-68 */
        if (e->series >= 0) e->Tw = series_update(&c->series[e->series], d->time);
        double kappa = co_Kappa(c, mat, Tc);
        double G[3] = {0.0, 0.0, 0.0}, gr[3];
        G[dir - 1] = (Tc - e->Tw) * modfm3;
        if (Fm <= 2) {
          G[1] = (TT[CX(B, Im, Jm - 1, Km)] - TT[CX(B, Im, Jm + 1, Km)]) / 2.0;
          G[2] = (TT[CX(B, Im, Jm, Km - 1)] - TT[CX(B, Im, Jm, Km + 1)]) / 2.0;
        } else if (Fm <= 4) {
          G[0] = (TT[CX(B, Im - 1, Jm, Km)] - TT[CX(B, Im + 1, Jm, Km)]) / 2.0;
          G[2] = (TT[CX(B, Im, Jm, Km - 1)] - TT[CX(B, Im, Jm, Km + 1)]) / 2.0;
        } else {
          G[0] = (TT[CX(B, Im - 1, Jm, Km)] - TT[CX(B, Im + 1, Jm, Km)]) / 2.0;
          G[1] = (TT[CX(B, Im, Jm - 1, Km)] - TT[CX(B, Im, Jm + 1, Km)]) / 2.0;
        }
        if (c->ndir == 2) G[2] = 0.0;
        matvec_GM(G, M, gr);
        double Flux = Area * kappa * dot3(gr, Nrm);
        R[p] = R[p] - modfm2 * Flux;
      } break;
      case 303: {
        /* BC_Wall_HeatTransfer, Lib_BC_Fluxes_Wall_HeatTransfer.f90:12-52.
         * The reference passes bc%qw, which is never set for 303
         * (src/lib/io/IO_BC.f90:203-206); it is 0 here. */
        double qw = 0.0;
        double kappa = co_Kappa(c, mat, Tc);
        double mn = MEL(M, dir, 1) * Nrm[0] + MEL(M, dir, 2) * Nrm[1] + MEL(M, dir, 3) * Nrm[2];
        double H = 1.0 / (2.0 * kappa * mn) + 1.0 / e->hconv;
        double Hrad = e->hconv / (2.0 * kappa * mn) + 1.0;
        double Flux = Area * (1.0 / H * (e->Tref - Tc) + 1.0 / Hrad * qw);
        R[p] = R[p] - Flux;
      } break;
      case 304:
      case 305: {
        /* BC_Wall_Radiation(+Convection), Lib_BC_Fluxes_Wall_Radiation.f90:12-65,
         * Lib_BC_Fluxes_Wall_RadiationConvection.f90:12-65 */
        double kappa = co_Kappa(c, mat, Tc);
        double mn = MEL(M, dir, 1) * Nrm[0] + MEL(M, dir, 2) * Nrm[1] + MEL(M, dir, 3) * Nrm[2];
        double Tw = Tc, qw = 0.0, err = 1.0e6, Flux = 0.0;
        double Tr2 = e->Tref * e->Tref, Tr4 = Tr2 * Tr2;
        int it = 0;
        while (err > 1.0e-3) {
          double Tw2 = Tw * Tw;
          double K = SIGMA_SB * e->eps * (Tw2 * Tw2 - Tr4) / (Tw - e->Tref);
          double H = (e->type == 304) ? 1.0 / (2.0 * kappa * mn) + 1.0 / K
                                      : 1.0 / (2.0 * kappa * mn) + 1.0 / (K + e->hconv);
          Flux = Area * (1.0 / H * (e->Tref - Tc));
          double q_wall = K * (e->Tref - Tw);
          err = fabs(qw - q_wall);
          qw = q_wall;
          Tw = e->Tref - 1.0 / (H * K) * (e->Tref - Tc);
          if (++it > 100000 || !(err == err)) {
            set_err(c, OB_ERR_STATE, e->b, Im, Jm, Km);
            break;
          }
        }
        R[p] = R[p] - Flux;
      } break;
      case 103:
        /* coupling flux slot, Mod_BC_Fluxes.f90:76-77 (ext_flux = 0 here) */
        R[p] = R[p] + 0.0;
        break;
      default:
        break;
    }
  }
}

/* Extract_Properties_From_BC (src/lib/io/IO_Wall.f90:131-239): the wall
 * temperature and heat flux (Ovar, Compute_Wall_Properties,
 * Lib_BC_Fluxes.f90:48-58) that the wall routines 301-305 return at the
 * current state; entries of other types get 0.  Boundary values of 301/302
 * are the ones the last BC_Fluxes used (their series at the level time). */
static void wall_props(ob_ctx *c, ob_dom *d, ob_bc *e, double *Tw_out, double *qw_out) {
  *Tw_out = 0.0;
  *qw_out = 0.0;
  if (e->type < 301 || e->type > 305) return;
  ob_blk *B = &d->blk[e->b - 1];
  int Im = e->i, Jm = e->j, Km = e->k, Fm = e->f;
  int dir, fi, fj, fk;
  face_index(Fm, Im, Jm, Km, &dir, &fi, &fj, &fk);
  long f = FX(B, dir - 1, fi, fj, fk);
  const double *Nrm = B->fN[dir - 1] + 3 * f;
  const double *M = Mc(B, Im, Jm, Km);
  const double *TT = B->T;
  long p = CX(B, Im, Jm, Km);
  double Tc = TT[p];
  double mat = B->matID[p];
  int modfm = Fm % 2, modfm2 = 1 - 2 * modfm, modfm3 = 2 * (2 * modfm - 1);
  double kappa = co_Kappa(c, mat, Tc);
  double mn = MEL(M, dir, 1) * Nrm[0] + MEL(M, dir, 2) * Nrm[1] + MEL(M, dir, 3) * Nrm[2];
  switch (e->type) {
    case 301: {
      /* Lib_BC_Fluxes_Wall_Heat.f90:37-52: T_wall = T - (G M)(Dir) */
      double G[3] = {0.0, 0.0, 0.0}, gr[3];
      G[dir - 1] = -e->qw / (2.0 * kappa * mn) * modfm2;
      if (c->ndir == 2) G[2] = 0.0;
      matvec_GM(G, M, gr);
      *Tw_out = Tc - gr[dir - 1];
      *qw_out = e->qw;
    } break;
    case 302: {
      /* Lib_BC_Fluxes_Wall_Temperature.f90:37-66: q_wall = kappa (G M) . n */
      double G[3] = {0.0, 0.0, 0.0}, gr[3];
      G[dir - 1] = (Tc - e->Tw) * modfm3;
      if (Fm <= 2) {
        G[1] = (TT[CX(B, Im, Jm - 1, Km)] - TT[CX(B, Im, Jm + 1, Km)]) / 2.0;
        G[2] = (TT[CX(B, Im, Jm, Km - 1)] - TT[CX(B, Im, Jm, Km + 1)]) / 2.0;
      } else if (Fm <= 4) {
        G[0] = (TT[CX(B, Im - 1, Jm, Km)] - TT[CX(B, Im + 1, Jm, Km)]) / 2.0;
        G[2] = (TT[CX(B, Im, Jm, Km - 1)] - TT[CX(B, Im, Jm, Km + 1)]) / 2.0;
      } else {
        G[0] = (TT[CX(B, Im - 1, Jm, Km)] - TT[CX(B, Im + 1, Jm, Km)]) / 2.0;
        G[1] = (TT[CX(B, Im, Jm - 1, Km)] - TT[CX(B, Im, Jm + 1, Km)]) / 2.0;
      }
      if (c->ndir == 2) G[2] = 0.0;
      matvec_GM(G, M, gr);
      *Tw_out = e->Tw;
      *qw_out = kappa * dot3(gr, Nrm);
    } break;
    case 303: {
      /* Lib_BC_Fluxes_Wall_HeatTransfer.f90:44-46 (qw unset for 303: 0) */
      double qw = 0.0;
      double Tw = (e->hconv * e->Tref + Tc * 2.0 * kappa * mn + qw) / (e->hconv + 2.0 * kappa * mn);
      *Tw_out = Tw;
      *qw_out = e->hconv * (e->Tref - Tw) + qw;
    } break;
    default: {
      /* 304 / 305: the fixed point of Lib_BC_Fluxes_Wall_Radiation.f90:41-57
       * (RadiationConvection: K + hconv in H); its last T_wall and q_wall */
      double Tw = Tc, qw = 0.0, err = 1.0e6, q_wall = 0.0;
      double Tr2 = e->Tref * e->Tref, Tr4 = Tr2 * Tr2;
      int it = 0;
      while (err > 1.0e-3) {
        double Tw2 = Tw * Tw;
        double K = SIGMA_SB * e->eps * (Tw2 * Tw2 - Tr4) / (Tw - e->Tref);
        double H = (e->type == 304) ? 1.0 / (2.0 * kappa * mn) + 1.0 / K
                                    : 1.0 / (2.0 * kappa * mn) + 1.0 / (K + e->hconv);
        q_wall = K * (e->Tref - Tw);
        err = fabs(qw - q_wall);
        qw = q_wall;
        Tw = e->Tref - 1.0 / (H * K) * (e->Tref - Tc);
        if (++it > 100000 || !(err == err)) {
          set_err(c, OB_ERR_STATE, e->b, Im, Jm, Km);
          break;
        }
      }
      *Tw_out = Tw;
      *qw_out = q_wall;
    } break;
  }
}

/* ------------------------------------------------------------------ */
/* time step: src/lib/numerics/time/Mod_dt.f90                          */
/* ------------------------------------------------------------------ */
static void compute_dt(ob_ctx *c, ob_dom *d) {
  double dtglobal = d->dtglobal;
  for (int bb = 0; bb < d->nb; bb++) {
    ob_blk *b = &d->blk[bb];
#ifdef _OPENMP
#pragma omp parallel for collapse(2) reduction(min : dtglobal)
#endif
    for (int k = 1; k <= b->n[2]; k++)
      for (int j = 1; j <= b->n[1]; j++)
        for (int i = 1; i <= b->n[0]; i++) {
          long p = CX(b, i, j, k);
          double T = b->T[p], m = b->matID[p];
          const double *dl = b->dl + 3 * p;
          double kappa = co_Kappa(c, m, T), rho = co_Rho(c, m, T), cp = co_Cp(c, m, T);
          double alpha = kappa / rho / cp;
          double dtmin = 1e8, invL2 = 0.0;
          for (int q = 0; q < c->ndir; q++) invL2 = invL2 + 1.0 / (dl[q] * dl[q]);
          double dt = 0.5 * c->vnn / (alpha * invL2);
          dtmin = dt < dtmin ? dt : dtmin;
          double dtcell = dtmin;
          if (d->iter < c->rampa) dtcell = dtcell * d->iter / c->rampa;
          b->dt[IX(b, i, j, k)] = dtcell;
          dtglobal = dtcell < dtglobal ? dtcell : dtglobal;
        }
  }
  d->dtglobal = dtglobal;
}

static void set_global_dt(ob_dom *d) {
  for (int bb = 0; bb < d->nb; bb++) {
    ob_blk *b = &d->blk[bb];
    long n = (long)b->n[0] * b->n[1] * b->n[2];
    for (long q = 0; q < n; q++) b->dt[q] = d->dtglobal;
  }
}

/* ------------------------------------------------------------------ */
/* IRS: src/lib/numerics/time/explicit/Lib_IRS.f90                      */
/* ------------------------------------------------------------------ */
static void irs_ghost(ob_dom *d) {
  /* Ghost_Residual_Extrapolation, Lib_IRS.f90:137-187 */
  for (int n = 0; n < d->nbound; n++) {
    ob_bc *e = &d->bc[n];
    ob_blk *bm = &d->blk[e->b - 1];
    int ig = e->i - GUIDE[e->f - 1][0], jg = e->j - GUIDE[e->f - 1][1], kg = e->k - GUIDE[e->f - 1][2];
    if (e->type == 101) {
      ob_blk *bs = &d->blk[e->bs - 1];
      bm->R[CX(bm, ig, jg, kg)] = bs->R[CX(bs, e->is, e->js, e->ks)];
    } else {
      bm->R[CX(bm, ig, jg, kg)] = bm->R[CX(bm, e->i, e->j, e->k)];
    }
  }
}

static void irs_blk(ob_blk *b, int dir, double beta) {
  /* Residual_Smoothing_Blk, Lib_IRS.f90:43-134 (njacobi = 2) */
  int n1 = b->n[0], n2 = b->n[1], n3 = b->n[2];
#ifdef _OPENMP
#pragma omp parallel for collapse(2)
#endif
  for (int k = 0; k <= n3 + 1; k++)
    for (int j = 0; j <= n2 + 1; j++)
      for (int i = 0; i <= n1 + 1; i++) b->RS1[CX(b, i, j, k)] = b->R[CX(b, i, j, k)];
  int di = dir == 1, dj = dir == 2, dk = dir == 3;
  for (int jac = 0; jac < 2; jac++) {
#ifdef _OPENMP
#pragma omp parallel for collapse(2)
#endif
    for (int k = 1; k <= n3; k++)
      for (int j = 1; j <= n2; j++)
        for (int i = 1; i <= n1; i++) {
          double r = b->R[CX(b, i, j, k)];
          double rm = b->RS1[CX(b, i - di, j - dj, k - dk)];
          double rp = b->RS1[CX(b, i + di, j + dj, k + dk)];
          b->RS2[CX(b, i, j, k)] = (r + beta * (rm + rp)) / (1.0 + 2.0 * beta);
        }
#ifdef _OPENMP
#pragma omp parallel for collapse(2)
#endif
    for (int k = 1; k <= n3; k++)
      for (int j = 1; j <= n2; j++)
        for (int i = 1; i <= n1; i++) b->RS1[CX(b, i, j, k)] = b->RS2[CX(b, i, j, k)];
  }
#ifdef _OPENMP
#pragma omp parallel for collapse(2)
#endif
  for (int k = 1; k <= n3; k++)
    for (int j = 1; j <= n2; j++)
      for (int i = 1; i <= n1; i++) b->R[CX(b, i, j, k)] = b->RS1[CX(b, i, j, k)];
}

/* Line-tridiagonal IRS (no reference counterpart): the fixed point of the
 * Jacobi iteration of Residual_Smoothing_Blk, i.e. the exact solution of
 *   (1+2b) x_s - b (x_{s-1} + x_{s+1}) = r_s,  s = 1..n,
 * with x_0, x_{n+1} the IRS ghosts of Ghost_Residual_Extrapolation held
 * fixed, by the Thomas algorithm in this order:
 *   den_1 = 1+2b, c_1 = -b/den_1, d_1 = (r_1 + b x_0 [+ b x_{n+1} if n=1]) / den_1
 *   den_s = (1+2b) + b c_{s-1}, c_s = -b/den_s,
 *   d_s = (r_s [+ b x_{n+1} if s=n] + b d_{s-1}) / den_s
 *   x_n = d_n, x_s = d_s - c_s x_{s+1}. */
static void irs_blk_tridiag(ob_blk *b, int dir, double beta) {
  int n1 = b->n[0], n2 = b->n[1], n3 = b->n[2];
  int di = dir == 1, dj = dir == 2, dk = dir == 3;
  int n = dir == 1 ? n1 : (dir == 2 ? n2 : n3);
  const double bb = 1.0 + 2.0 * beta;
  double *cp = (double *)malloc(sizeof(double) * (n + 1));
  double *dp = (double *)malloc(sizeof(double) * (n + 1));
  /* lines: every (i,j,k) with the dir index 1 */
  for (int k = 1; k <= (dk ? 1 : n3); k++)
    for (int j = 1; j <= (dj ? 1 : n2); j++)
      for (int i = 1; i <= (di ? 1 : n1); i++) {
        const double glo = b->R[CX(b, i - di, j - dj, k - dk)];
        const double ghi = b->R[CX(b, i + di * n, j + dj * n, k + dk * n)];
        for (int s = 1; s <= n; s++) {
          double r = b->R[CX(b, i + di * (s - 1), j + dj * (s - 1), k + dk * (s - 1))];
          double den;
          if (s == 1) {
            den = bb;
            r = r + beta * glo;
          } else {
            den = bb + beta * cp[s - 1];
          }
          if (s == n) r = r + beta * ghi;
          cp[s] = -beta / den;
          dp[s] = s == 1 ? r / den : (r + beta * dp[s - 1]) / den;
        }
        double x = dp[n];
        b->RS1[CX(b, i + di * (n - 1), j + dj * (n - 1), k + dk * (n - 1))] = x;
        for (int s = n - 1; s >= 1; s--) {
          x = dp[s] - cp[s] * x;
          b->RS1[CX(b, i + di * (s - 1), j + dj * (s - 1), k + dk * (s - 1))] = x;
        }
      }
  for (int k = 1; k <= n3; k++)
    for (int j = 1; j <= n2; j++)
      for (int i = 1; i <= n1; i++) b->R[CX(b, i, j, k)] = b->RS1[CX(b, i, j, k)];
  free(cp);
  free(dp);
}

static void residual_smoothing(ob_ctx *c, ob_dom *d) {
  /* Residual_Smoothing, Lib_IRS.f90:14-40 */
  for (int dir = 1; dir <= c->ndir; dir++) {
    irs_ghost(d);
    for (int bb = 0; bb < d->nb; bb++) {
      if (c->irs_mode == 1) irs_blk_tridiag(&d->blk[bb], dir, c->beta);
      else irs_blk(&d->blk[bb], dir, c->beta);
    }
  }
}

/* ------------------------------------------------------------------ */
/* state update: src/lib/numerics/time/explicit/Lib_Newstate.f90         */
/* ------------------------------------------------------------------ */
static inline double rk_stage(ob_ctx *c, int irk, double state, double old, double res) {
  /* RK_stage, Lib_RK.f90:15-26 */
  if (c->n_rk == 4) return old + res * RK4A[irk - 1];
  double r_ = (state - old + res) * RKC[c->n_rk - 1][irk - 1];
  return old + r_;
}
static int check_state(ob_ctx *c, double T, int b, int i, int j, int k) {
  /* Check_And_Fix_State, Lib_Newstate.f90:278-290 */
  if (isnan(T) || T < 0.0) {
    set_err(c, OB_ERR_STATE, b, i, j, k);
    return 1;
  }
  return 0;
}

static void newstate(ob_ctx *c, ob_dom *d, int irk) {
  if (c->prim) {
    /* Newstate_Primitive, Lib_Newstate.f90:154-272 */
    for (int bb = 0; bb < d->nb; bb++) {
      ob_blk *b = &d->blk[bb];
#ifdef _OPENMP
#pragma omp parallel for collapse(2)
#endif
      for (int k = 1; k <= b->n[2]; k++)
        for (int j = 1; j <= b->n[1]; j++)
          for (int i = 1; i <= b->n[0]; i++) {
            long p = CX(b, i, j, k);
            double T = b->T[p], m = b->matID[p];
            double rho = co_Rho(c, m, T), cp = co_Cp(c, m, T);
            double drhodT = co_slope(c, c->rho, m, T), dcpdT = co_slope(c, c->cp, m, T);
            /* Compute_Prim_Residual, Lib_Newstate.f90:241-256 */
            b->R[p] = (-b->R[p] / b->vol[p] + b->qvol[p]) * b->dt[IX(b, i, j, k)] /
                      (T * cp * drhodT + rho * cp + T * rho * dcpdT);
            if (!c->irs) {
              b->T[p] = rk_stage(c, irk, b->T[p], b->TO[p], b->R[p]);
              check_state(c, b->T[p], bb + 1, i, j, k);
            }
          }
    }
    if (c->irs) {
      residual_smoothing(c, d);
      for (int bb = 0; bb < d->nb; bb++) {
        ob_blk *b = &d->blk[bb];
#ifdef _OPENMP
#pragma omp parallel for collapse(2)
#endif
        for (int k = 1; k <= b->n[2]; k++)
          for (int j = 1; j <= b->n[1]; j++)
            for (int i = 1; i <= b->n[0]; i++) {
              long p = CX(b, i, j, k);
              b->T[p] = rk_stage(c, irk, b->T[p], b->TO[p], b->R[p]);
              check_state(c, b->T[p], bb + 1, i, j, k);
            }
      }
    }
  } else {
    /* Newstate_Conservative, Lib_Newstate.f90:13-148 */
    for (int bb = 0; bb < d->nb; bb++) {
      ob_blk *b = &d->blk[bb];
#ifdef _OPENMP
#pragma omp parallel for collapse(2)
#endif
      for (int k = 1; k <= b->n[2]; k++)
        for (int j = 1; j <= b->n[1]; j++)
          for (int i = 1; i <= b->n[0]; i++) {
            long p = CX(b, i, j, k);
            b->R[p] = (-b->R[p] / b->vol[p] + b->qvol[p]) * b->dt[IX(b, i, j, k)];
            if (!c->irs) {
              double m = b->matID[p];
              double hO = co_H(c, m, b->TO[p]), h = co_H(c, m, b->T[p]);
              h = rk_stage(c, irk, h, hO, b->R[p]);
              b->T[p] = co_T(c, m, h);
              check_state(c, b->T[p], bb + 1, i, j, k);
            }
          }
    }
    if (c->irs) {
      residual_smoothing(c, d);
      for (int bb = 0; bb < d->nb; bb++) {
        ob_blk *b = &d->blk[bb];
#ifdef _OPENMP
#pragma omp parallel for collapse(2)
#endif
        for (int k = 1; k <= b->n[2]; k++)
          for (int j = 1; j <= b->n[1]; j++)
            for (int i = 1; i <= b->n[0]; i++) {
              long p = CX(b, i, j, k);
              double m = b->matID[p];
              double hO = co_H(c, m, b->TO[p]), h = co_H(c, m, b->T[p]);
              h = rk_stage(c, irk, h, hO, b->R[p]);
              b->T[p] = co_T(c, m, h);
              check_state(c, b->T[p], bb + 1, i, j, k);
            }
      }
    }
  }
}

/* Compute_Residual, src/lib/diagnostic/Mod_Diagnostic.f90:13-49 */
static double compute_residual(ob_dom *d) {
  double tot = 0.0;
  for (int bb = 0; bb < d->nb; bb++) {
    ob_blk *b = &d->blk[bb];
    double res = 0.0;
#ifdef _OPENMP
#pragma omp parallel for collapse(2) reduction(+ : res)
#endif
    for (int k = 1; k <= b->n[2]; k++)
      for (int j = 1; j <= b->n[1]; j++)
        for (int i = 1; i <= b->n[0]; i++) {
          long p = CX(b, i, j, k);
          double r = fabs(b->T[p] - b->TO[p]);
          res = res + r * r;
        }
    tot = tot + res;
  }
  return tot;
}

/* Explicit_Step, src/lib/numerics/time/explicit/Mod_Explicit.f90:10-62 */
int ob_step(ob_ctx *c, int level, double *res_sq) {
  if (!c || level < 1 || level > c->nlev || !c->setup_done) return OB_ERR_ARG;
  ob_dom *d = &c->dom[level - 1];
  d->iter = d->iter + 1;
  d->dtglobal = 1e5;
  compute_dt(c, d);
  if (c->time_accurate) {
    set_global_dt(d);
    d->time = d->time + d->dtglobal;
  }
  /* Copy_State, Mod_Explicit.f90:114-135 */
  for (int bb = 0; bb < d->nb; bb++) {
    ob_blk *b = &d->blk[bb];
#ifdef _OPENMP
#pragma omp parallel for collapse(2)
#endif
    for (int k = 1; k <= b->n[2]; k++)
      for (int j = 1; j <= b->n[1]; j++)
        for (int i = 1; i <= b->n[0]; i++) b->TO[CX(b, i, j, k)] = b->T[CX(b, i, j, k)];
  }
  for (int irk = 1; irk <= c->n_rk; irk++) {
    ghost_fill(c, d);
    for (int bb = 0; bb < d->nb; bb++) fluxes_blk(c, &d->blk[bb]);
    bc_fluxes(c, d);
    newstate(c, d, irk);
    if (c->err) return c->err;
  }
  if (res_sq) *res_sq = compute_residual(d);
  return c->err;
}

/* ------------------------------------------------------------------ */
/* multigrid: src/lib/numerics/multigrid/Lib_Multigrid.f90              */
/* ------------------------------------------------------------------ */
/* fine2coarse_prim, Lib_Multigrid.f90:38-125 */
static void fine2coarse(ob_ctx *c, ob_blk *f, ob_blk *cb) {
  int fn3 = f->n[2];
  long nf = (long)f->n[0] * f->n[1] * f->n[2];
  double *fC = (double *)malloc(sizeof(double) * nf);
  for (int k = 1; k <= f->n[2]; k++)
    for (int j = 1; j <= f->n[1]; j++)
      for (int i = 1; i <= f->n[0]; i++) {
        long p = CX(f, i, j, k);
        fC[IX(f, i, j, k)] = co_H(c, f->matID[p], f->T[p]);
      }
#define FC(i, j, k) fC[IX(f, i, j, k)]
#define FV(i, j, k) f->vol[CX(f, i, j, k)]
#define FQ(i, j, k) f->qvol[CX(f, i, j, k)]
  for (int k = 1; k <= cb->n[2]; k++)
    for (int j = 1; j <= cb->n[1]; j++)
      for (int i = 1; i <= cb->n[0]; i++) {
        int i2 = 2 * i, j2 = 2 * j, k2 = 2 * k, i2d = i2 - 1, j2d = j2 - 1, k2d = k2 - 1;
        double cc, cq, tv;
        long pc = CX(cb, i, j, k);
        if (fn3 == 1) {
          cc = FC(i2d, j2d, 1) * FV(i2d, j2d, 1) + FC(i2, j2d, 1) * FV(i2, j2d, 1) +
               FC(i2d, j2, 1) * FV(i2d, j2, 1) + FC(i2, j2, 1) * FV(i2, j2, 1);
          cq = FQ(i2d, j2d, 1) * FV(i2d, j2d, 1) + FQ(i2, j2d, 1) * FV(i2, j2d, 1) +
               FQ(i2d, j2, 1) * FV(i2d, j2, 1) + FQ(i2, j2, 1) * FV(i2, j2, 1);
          tv = FV(i2d, j2d, 1) + FV(i2, j2d, 1) + FV(i2d, j2, 1) + FV(i2, j2, 1);
          k2 = 1;
        } else {
          cc = FC(i2d, j2d, k2d) * FV(i2d, j2d, k2d) + FC(i2, j2d, k2d) * FV(i2, j2d, k2d) +
               FC(i2d, j2, k2d) * FV(i2d, j2, k2d) + FC(i2, j2, k2d) * FV(i2, j2, k2d) +
               FC(i2d, j2d, k2) * FV(i2d, j2d, k2) + FC(i2, j2d, k2) * FV(i2, j2d, k2) +
               FC(i2d, j2, k2) * FV(i2d, j2, k2) + FC(i2, j2, k2) * FV(i2, j2, k2);
          cq = FQ(i2d, j2d, k2d) * FV(i2d, j2d, k2d) + FQ(i2, j2d, k2d) * FV(i2, j2d, k2d) +
               FQ(i2d, j2, k2d) * FV(i2d, j2, k2d) + FQ(i2, j2, k2d) * FV(i2, j2, k2d) +
               FQ(i2d, j2d, k2) * FV(i2d, j2d, k2) + FQ(i2, j2d, k2) * FV(i2, j2d, k2) +
               FQ(i2d, j2, k2) * FV(i2d, j2, k2) + FQ(i2, j2, k2) * FV(i2, j2, k2);
          tv = FV(i2d, j2d, k2d) + FV(i2, j2d, k2d) + FV(i2d, j2, k2d) + FV(i2, j2, k2d) +
               FV(i2d, j2d, k2) + FV(i2, j2d, k2) + FV(i2d, j2, k2) + FV(i2, j2, k2);
        }
        cb->qvol[pc] = cq / tv;
        /* cmatID(i,j,k) = fmatID(2i,2j,2k): in 2-D k2=2k=2 is the k-ghost of
         * the single fine plane (the mirror of plane 1) */
        cb->matID[pc] = f->matID[CX(f, i2, j2, 2 * k)];
        cb->R[pc] = cc; /* cCons held in R until the division below */
      }
  for (int k = 1; k <= cb->n[2]; k++)
    for (int j = 1; j <= cb->n[1]; j++)
      for (int i = 1; i <= cb->n[0]; i++) {
        long pc = CX(cb, i, j, k);
        double cc = cb->R[pc] / cb->vol[pc];
        cb->T[pc] = co_T(c, cb->matID[pc], cc);
        cb->R[pc] = 0.0;
      }
#undef FC
#undef FV
#undef FQ
  free(fC);
}

/* coarse2fine_prim, Lib_Multigrid.f90:128-220 */
static void coarse2fine(ob_blk *f, ob_blk *cb) {
  double co[8];
  if (f->n[2] == 1) {
    double a1 = 9.0 / 16.0, a2 = 3.0 / 16.0, a3 = 1.0 / 16.0, a4 = 0.0;
    double t[8] = {a1, a2, a2, a4, a3, a4, a4, a4};
    memcpy(co, t, sizeof(co));
  } else {
    double a1 = 27.0 / 64.0, a2 = 9.0 / 64.0, a3 = 3.0 / 64.0, a4 = 1.0 / 64.0;
    double t[8] = {a1, a2, a2, a2, a3, a3, a3, a4};
    memcpy(co, t, sizeof(co));
  }
  static const int MASK[8][3] = {{1, 2, 3}, {4, 2, 3}, {1, 5, 3}, {4, 5, 3},
                                 {1, 2, 6}, {4, 2, 6}, {1, 5, 6}, {4, 5, 6}};
#define CP(i, j, k) cb->T[CX(cb, i, j, k)]
  for (int k = 1; k <= cb->n[2]; k++)
    for (int j = 1; j <= cb->n[1]; j++)
      for (int i = 1; i <= cb->n[0]; i++) {
        int i2 = 2 * i, j2 = 2 * j, k2 = 2 * k, i2d = i2 - 1, j2d = j2 - 1, k2d = k2 - 1;
        int id[6];
        id[0] = i - 1 > 1 ? i - 1 : 1;
        id[1] = j - 1 > 1 ? j - 1 : 1;
        id[2] = k - 1 > 1 ? k - 1 : 1;
        id[3] = i + 1 < cb->n[0] ? i + 1 : cb->n[0];
        id[4] = j + 1 < cb->n[1] ? j + 1 : cb->n[1];
        id[5] = k + 1 < cb->n[2] ? k + 1 : cb->n[2];
        if (f->n[2] == 1) { k2 = 1; k2d = 1; }
        int counter = 0;
        for (int kk = k2d; kk <= k2; kk++)
          for (int jj = j2d; jj <= j2; jj++)
            for (int ii = i2d; ii <= i2; ii++) {
              const int *mk = MASK[counter];
              int a = id[mk[0] - 1], bq = id[mk[1] - 1], cq = id[mk[2] - 1];
              double v = co[0] * CP(i, j, k) + co[1] * CP(a, j, k) + co[2] * CP(i, bq, k) +
                         co[3] * CP(i, j, cq) + co[4] * CP(a, bq, k) + co[5] * CP(a, j, cq) +
                         co[6] * CP(i, bq, cq) + co[7] * CP(a, bq, cq);
              f->T[CX(f, ii, jj, kk)] = v;
              counter++;
            }
      }
#undef CP
}

int ob_restrict(ob_ctx *c, int fine_level) {
  if (!c || fine_level < 1 || fine_level >= c->nlev) return OB_ERR_ARG;
  ob_dom *F = &c->dom[fine_level - 1], *C = &c->dom[fine_level];
  for (int b = 0; b < F->nb; b++) fine2coarse(c, &F->blk[b], &C->blk[b]);
  return c->err;
}
int ob_prolong(ob_ctx *c, int fine_level) {
  if (!c || fine_level < 1 || fine_level >= c->nlev) return OB_ERR_ARG;
  ob_dom *F = &c->dom[fine_level - 1], *C = &c->dom[fine_level];
  for (int b = 0; b < F->nb; b++) coarse2fine(&F->blk[b], &C->blk[b]);
  return c->err;
}

/* ------------------------------------------------------------------ */
/* construction                                                        */
/* ------------------------------------------------------------------ */
static int alloc_blk(ob_blk *b, int ni, int nj, int nk) {
  b->n[0] = ni; b->n[1] = nj; b->n[2] = nk;
  b->s1 = ni + 2 * GC;
  b->s2 = b->s1 * (nj + 2 * GC);
  b->npad = b->s2 * (nk + 2 * GC);
  long ni_ = (long)ni * nj * nk;
  b->T = calloc(b->npad, sizeof(double));
  b->TO = calloc(b->npad, sizeof(double));
  b->R = calloc(b->npad, sizeof(double));
  b->RS1 = calloc(b->npad, sizeof(double));
  b->RS2 = calloc(b->npad, sizeof(double));
  b->qvol = calloc(b->npad, sizeof(double));
  b->matID = calloc(b->npad, sizeof(double));
  b->vol = calloc(b->npad, sizeof(double));
  b->M = calloc(9 * b->npad, sizeof(double));
  b->dl = calloc(3 * b->npad, sizeof(double));
  b->dt = calloc(ni_, sizeof(double));
  b->node = calloc(3L * (ni + 1) * (nj + 1) * (nk + 1), sizeof(double));
  for (int d = 0; d < 3; d++) {
    b->fN[d] = calloc(3 * nfaces(b, d), sizeof(double));
    b->fA[d] = calloc(nfaces(b, d), sizeof(double));
  }
  return (b->T && b->M && b->node) ? 0 : OB_ERR_ARG;
}
static void free_blk(ob_blk *b) {
  free(b->T); free(b->TO); free(b->R); free(b->RS1); free(b->RS2); free(b->qvol);
  free(b->matID); free(b->vol); free(b->M); free(b->dl); free(b->dt); free(b->node);
  for (int d = 0; d < 3; d++) { free(b->fN[d]); free(b->fA[d]); }
}

ob_ctx *ob_create(int nlevels, int nb) {
  if (nlevels < 1 || nb < 1) return NULL;
  ob_ctx *c = calloc(1, sizeof(ob_ctx));
  c->nlev = nlevels;
  c->nb = nb;
  c->dom = calloc(nlevels, sizeof(ob_dom));
  for (int m = 0; m < nlevels; m++) {
    c->dom[m].nb = nb;
    c->dom[m].blk = calloc(nb, sizeof(ob_blk));
  }
  c->n_rk = 1;
  c->vnn = 0.3;
  c->prim = 0;
  return c;
}

void ob_destroy(ob_ctx *c) {
  if (!c) return;
  for (int m = 0; m < c->nlev; m++) {
    for (int b = 0; b < c->nb; b++)
      if (c->dom[m].blk[b].T) free_blk(&c->dom[m].blk[b]);
    free(c->dom[m].blk);
    for (int n = 0; n < c->dom[m].nbound; n++) { free(c->dom[m].bc[n].cdon); free(c->dom[m].bc[n].cvf); }
    free(c->dom[m].bc);
    free(c->dom[m].n_bf);
  }
  free(c->dom);
  for (int s = 0; s < c->nseries; s++) { free(c->series[s].t); free(c->series[s].v); }
  free(c->series);
  free(c->cp); free(c->rho); free(c->kap); free(c->h);
  free(c);
}

int ob_set_table(ob_ctx *c, int nmat, int Ti1, int nT, const double *cp, const double *rho,
                 const double *kap, const double *h) {
  if (!c || nmat < 1 || nT < 2) return OB_ERR_ARG;
  c->nmat = nmat; c->Ti1 = Ti1; c->nT = nT; c->Ti2 = Ti1 + nT - 1;
  size_t n = (size_t)nmat * nT;
  c->cp = malloc(n * sizeof(double)); memcpy(c->cp, cp, n * sizeof(double));
  c->rho = malloc(n * sizeof(double)); memcpy(c->rho, rho, n * sizeof(double));
  c->kap = malloc(n * sizeof(double)); memcpy(c->kap, kap, n * sizeof(double));
  c->h = malloc(n * sizeof(double)); memcpy(c->h, h, n * sizeof(double));
  for (int m = 0; m < nmat; m++)
    for (int t = 1; t < nT; t++)
      if (c->h[(long)m * nT + t] < c->h[(long)m * nT + t - 1]) return OB_ERR_TABLE;
  return 0;
}

int ob_set_block(ob_ctx *c, int b, int ni, int nj, int nk, const double *nodes,
                 const double *T, const double *matID, const double *qvol) {
  if (!c || b < 1 || b > c->nb || ni < 1 || nj < 1 || nk < 1) return OB_ERR_ARG;
  ob_blk *B = &c->dom[0].blk[b - 1];
  if (B->T) free_blk(B);
  if (alloc_blk(B, ni, nj, nk)) return OB_ERR_ARG;
  memcpy(B->node, nodes, sizeof(double) * 3L * (ni + 1) * (nj + 1) * (nk + 1));
  for (int k = 1; k <= nk; k++)
    for (int j = 1; j <= nj; j++)
      for (int i = 1; i <= ni; i++) {
        long p = CX(B, i, j, k), q = IX(B, i, j, k);
        B->T[p] = T[q];
        B->matID[p] = matID[q];
        B->qvol[p] = qvol ? qvol[q] : 0.0;
      }
  return 0;
}

int ob_add_series(ob_ctx *c, int periodic, int n, const double *t, const double *v) {
  if (!c || n < 1) return -1;
  c->series = realloc(c->series, sizeof(ob_series) * (c->nseries + 1));
  ob_series *s = &c->series[c->nseries];
  s->periodic = periodic; s->n = n;
  s->t = malloc(sizeof(double) * n); memcpy(s->t, t, sizeof(double) * n);
  s->v = malloc(sizeof(double) * n); memcpy(s->v, v, sizeof(double) * n);
  return c->nseries++;
}

int ob_set_bc(ob_ctx *c, int level, int nbound, const int *ib, const double *rb,
              const int *series) {
  if (!c || level < 1 || level > c->nlev || nbound < 0) return OB_ERR_ARG;
  ob_dom *d = &c->dom[level - 1];
  for (int n = 0; d->bc && n < d->nbound; n++) { free(d->bc[n].cdon); free(d->bc[n].cvf); }
  free(d->bc);
  free(d->n_bf);
  d->nbound = nbound;
  d->bc = calloc(nbound > 0 ? nbound : 1, sizeof(ob_bc));
  d->n_bf = calloc(6 * d->nb, sizeof(int));
  for (int n = 0; n < nbound; n++) {
    const int *q = ib + OB_BC_NI * n;
    const double *r = rb + OB_BC_NR * n;
    ob_bc *e = &d->bc[n];
    e->b = q[0]; e->i = q[1]; e->j = q[2]; e->k = q[3]; e->f = q[4]; e->type = q[5];
    e->bs = q[6]; e->is = q[7]; e->js = q[8]; e->ks = q[9]; e->fs = q[10];
    e->d11 = q[11]; e->d12 = q[12]; e->d21 = q[13]; e->d22 = q[14];
    e->qw = r[0]; e->Tw = r[1]; e->hconv = r[2]; e->qrad = r[3]; e->Tref = r[4]; e->eps = r[5];
    e->series = series ? series[n] : -1;
    if (e->b < 1 || e->b > d->nb || e->f < 1 || e->f > 6) return OB_ERR_ARG;
    if (e->type == 101 && (e->bs < 1 || e->bs > d->nb || e->fs < 1 || e->fs > 6)) return OB_ERR_ARG;
    d->n_bf[(e->b - 1) * 6 + (e->f - 1)]++;
  }
  return 0;
}

int ob_set_chimera(ob_ctx *c, int level, int n, const int *entry, const int *n1, const int *n2,
                   const int *donors, const double *vf) {
  if (!c || level < 1 || level > c->nlev || n < 0) return OB_ERR_ARG;
  ob_dom *d = &c->dom[level - 1];
  long off = 0;
  for (int q = 0; q < n; q++) {
    if (entry[q] < 0 || entry[q] >= d->nbound || n1[q] < 1 || n2[q] < 0) return OB_ERR_ARG;
    ob_bc *e = &d->bc[entry[q]];
    if (e->type != 102) return OB_ERR_ARG;
    int m = n1[q] + n2[q];
    free(e->cdon); free(e->cvf);
    e->cdon = malloc(4 * (size_t)m * sizeof(int));
    e->cvf = malloc((size_t)m * sizeof(double));
    memcpy(e->cdon, donors + 4 * off, 4 * (size_t)m * sizeof(int));
    memcpy(e->cvf, vf + off, (size_t)m * sizeof(double));
    e->nch1 = n1[q]; e->nch = m;
    for (int t = 0; t < m; t++) {
      const int *dn = e->cdon + 4 * t;
      if (dn[0] < 1 || dn[0] > d->nb) return OB_ERR_ARG;
      ob_blk *bs = &d->blk[dn[0] - 1];
      if (dn[1] < 1 || dn[1] > bs->n[0] || dn[2] < 1 || dn[2] > bs->n[1] || dn[3] < 1 || dn[3] > bs->n[2])
        return OB_ERR_ARG;
    }
    off += m;
  }
  return 0;
}

int ob_set_params(ob_ctx *c, int n_rk, double vnn, int rampa_iter, int time_accurate,
                  int prim, double irs_beta) {
  if (!c || n_rk < 1 || n_rk > 4) return OB_ERR_ARG;
  c->n_rk = n_rk; c->vnn = vnn; c->rampa = rampa_iter; c->time_accurate = time_accurate;
  c->prim = prim; c->beta = irs_beta; c->irs = irs_beta > 0.0;
  return 0;
}

int ob_set_irs_mode(ob_ctx *c, int mode) {
  if (!c || mode < 0 || mode > 1) return OB_ERR_ARG;
  c->irs_mode = mode;
  return 0;
}

int ob_setup(ob_ctx *c, double initial_time) {
  if (!c || !c->kap) return OB_ERR_ARG;
  ob_dom *d1 = &c->dom[0];
  for (int b = 0; b < c->nb; b++)
    if (!d1->blk[b].T) return OB_ERR_ARG;
  /* Check_Mesh_Type, Lib_Metrics.f90:13-66 (ndir from block 1) */
  ob_blk *b1 = &d1->blk[0];
  if (b1->n[2] > 1) c->ndir = 3;
  else if (b1->n[2] == 1 && b1->n[1] == 1) c->ndir = 1;
  else c->ndir = 2;
  /* Setup_Multigrid / Coarse_Grid, Mod_Multigrid.f90:10-56, Lib_Multigrid.f90:223-265 */
  for (int m = 1; m < c->nlev; m++) {
    for (int b = 0; b < c->nb; b++) {
      ob_blk *F = &c->dom[m - 1].blk[b], *C = &c->dom[m].blk[b];
      int cn[3] = {F->n[0] / 2, F->n[1] / 2, F->n[2] / 2};
      if (cn[2] < 1) cn[2] = 1;
      int rap = 1 << m; /* Check_Multigrid, Lib_Multigrid.f90:11-35 (on level 1 dims) */
      ob_blk *B1 = &c->dom[0].blk[b];
      int chk = 0;
      for (int q = 0; q < 3; q++) {
        if (B1->n[q] % (1 << (c->nlev - 1)) == 0) chk++;
        if (q == 2 && B1->n[q] == 1) chk++;
      }
      (void)rap;
      if (chk < 3) return OB_ERR_ARG;
      if (C->T) free_blk(C);
      alloc_blk(C, cn[0], cn[1], cn[2]);
      int step = 2 - F->n[2] % 2;
      for (int k = 0; k <= F->n[2]; k += step)
        for (int j = 0; j <= F->n[1]; j += 2)
          for (int i = 0; i <= F->n[0]; i += 2) {
            int k2 = (F->n[2] == 1) ? k : k / 2;
            memcpy(C->node + 3 * NX(C, i / 2, j / 2, k2), F->node + 3 * NX(F, i, j, k),
                   3 * sizeof(double));
          }
    }
  }
  for (int m = 0; m < c->nlev; m++) {
    ob_dom *d = &c->dom[m];
    if (!d->bc) return OB_ERR_ARG;
    fill_matIDg(d);
    for (int b = 0; b < c->nb; b++)
      if (setup_block_metrics(c, &d->blk[b])) return c->err;
    for (int n = 0; n < d->nbound; n++) {
      ob_bc *e = &d->bc[n];
      if (e->type == 101) {
        int r = connect_metrics(c, e, &d->blk[e->b - 1], &d->blk[e->bs - 1]);
        if (r) return r;
      } else if (e->type == 102) {
        /* donor lists are given for the finest level only */
        if (m > 0 || e->nch1 < 1) {
          set_err(c, OB_ERR_UNSUPPORTED, e->b, e->i, e->j, e->k);
          return OB_ERR_UNSUPPORTED;
        }
        int r = extrap_metric(c, e, &d->blk[e->b - 1]);
        if (r) return r;
      }
    }
    d->iter = 0;
    d->time = initial_time;
  }
  c->setup_done = 1;
  /* FUSS_setup tail, src/lib/driver/Wrap_Setup.f90:145-153 */
  ghost_fill(c, &c->dom[0]);
  for (int m = 1; m < c->nlev; m++) {
    for (int b = 0; b < c->nb; b++) fine2coarse(c, &c->dom[m - 1].blk[b], &c->dom[m].blk[b]);
    ghost_fill(c, &c->dom[m]);
  }
  /* fill_matIDg runs before the metrics in the reference (Wrap_Setup.f90:70-76)
   * on ghost matIDs that the first ghost fill sets; the first ghost fill also
   * sets matIDg for connections, so redo it for faithfulness. */
  for (int m = 0; m < c->nlev; m++) fill_matIDg(&c->dom[m]);
  return c->err;
}

/* ------------------------------------------------------------------ */
/* driver: FUSS_solve + Explicit_Step control + FUSS_postprocess MG switch */
/* ------------------------------------------------------------------ */
int ob_run(ob_ctx *c, const int *itermax, double res_threshold, double time_threshold,
           int max_iters, int hist_cap, int *n_iter, int *hist_iter, double *hist_time,
           double *hist_res, int nprobe, const int *probes, double *probe_out) {
  if (!c || !c->setup_done) return OB_ERR_ARG;
  int level = c->nlev, iter_from_call = 0, iter_general = 0, nit = 0;
  double residuotot = 0.0;
  for (int m = 0; m < c->nlev; m++) c->dom[m].itermax = itermax[m];
  int nrow_probe = 0;
  while (nit < max_iters) {
    ob_dom *d = &c->dom[level - 1];
    iter_from_call++;
    iter_general++;
    double rsq = 0.0;
    int st = ob_step(c, level, &rsq);
    nit++;
    if (st) {
      *n_iter = nit;
      return st;
    }
    /* res_diter = 1: residual every iteration (Mod_Explicit.f90:65-78) */
    residuotot = sqrt(rsq);
    if (nit <= hist_cap) {
      hist_iter[nit - 1] = iter_general;
      hist_time[nit - 1] = d->time;
      hist_res[nit - 1] = residuotot;
    }
    if (level == 1 && nprobe > 0 && probe_out) {
      for (int p = 0; p < nprobe; p++) {
        const int *q = probes + 4 * p;
        ob_blk *B = &c->dom[0].blk[q[0] - 1];
        probe_out[(long)nrow_probe * (nprobe + 1) + 0] = d->time;
        probe_out[(long)nrow_probe * (nprobe + 1) + 1 + p] = B->T[CX(B, q[1], q[2], q[3])];
      }
      nrow_probe++;
    }
    int stop = (iter_from_call >= d->itermax) || (residuotot <= res_threshold) ||
               (d->time >= time_threshold);
    if (level == 1) {
      if (stop) break;
    } else if (stop) {
      /* change_MG: Prolongation(fine=level-1, coarse=level), Wrap_Solve.f90:27-29,
       * then MG_level -= 1, iter_from_call = 0 (Wrap_Postprocess.f90:183-187) */
      ob_prolong(c, level - 1);
      level--;
      iter_from_call = 0;
    }
  }
  *n_iter = nit;
  return c->err;
}

/* ------------------------------------------------------------------ */
/* accessors                                                           */
/* ------------------------------------------------------------------ */
int ob_get(ob_ctx *c, int level, int b, const char *name, double *out) {
  if (!c || level < 1 || level > c->nlev || b < 1 || b > c->nb) return OB_ERR_ARG;
  ob_blk *B = &c->dom[level - 1].blk[b - 1];
  if (!strcmp(name, "Tpad")) {
    memcpy(out, B->T, sizeof(double) * B->npad);
    return 0;
  }
  if (!strcmp(name, "dt")) {
    memcpy(out, B->dt, sizeof(double) * B->n[0] * B->n[1] * B->n[2]);
    return 0;
  }
  double *src = NULL;
  if (!strcmp(name, "T")) src = B->T;
  else if (!strcmp(name, "TO")) src = B->TO;
  else if (!strcmp(name, "R")) src = B->R;
  else if (!strcmp(name, "vol")) src = B->vol;
  else if (!strcmp(name, "qvol")) src = B->qvol;
  else if (!strcmp(name, "matID")) src = B->matID;
  if (!src) return OB_ERR_ARG;
  for (int k = 1; k <= B->n[2]; k++)
    for (int j = 1; j <= B->n[1]; j++)
      for (int i = 1; i <= B->n[0]; i++) out[IX(B, i, j, k)] = src[CX(B, i, j, k)];
  return 0;
}

int ob_set(ob_ctx *c, int level, int b, const char *name, const double *in) {
  if (!c || level < 1 || level > c->nlev || b < 1 || b > c->nb) return OB_ERR_ARG;
  ob_blk *B = &c->dom[level - 1].blk[b - 1];
  double *dst = NULL;
  if (!strcmp(name, "T")) dst = B->T;
  else if (!strcmp(name, "qvol")) dst = B->qvol;
  if (!dst) return OB_ERR_ARG;
  for (int k = 1; k <= B->n[2]; k++)
    for (int j = 1; j <= B->n[1]; j++)
      for (int i = 1; i <= B->n[0]; i++) dst[CX(B, i, j, k)] = in[IX(B, i, j, k)];
  return 0;
}

int ob_dims(ob_ctx *c, int level, int b, int *dims) {
  if (!c || level < 1 || level > c->nlev || b < 1 || b > c->nb) return OB_ERR_ARG;
  ob_blk *B = &c->dom[level - 1].blk[b - 1];
  dims[0] = B->n[0]; dims[1] = B->n[1]; dims[2] = B->n[2];
  return 0;
}
int ob_ndir(ob_ctx *c) { return c ? c->ndir : 0; }
/* threads of the OpenMP build (1 for the scalar checker) */
int ob_threads(void) {
#ifdef _OPENMP
  return omp_get_max_threads();
#else
  return 1;
#endif
}
double ob_time(ob_ctx *c, int level) { return c->dom[level - 1].time; }
double ob_dtglobal(ob_ctx *c, int level) { return c->dom[level - 1].dtglobal; }
int ob_last_error(ob_ctx *c, int *bijk) {
  if (bijk) { bijk[0] = c->eb; bijk[1] = c->ei; bijk[2] = c->ej; bijk[3] = c->ek; }
  return c->err;
}

int ob_get_metrics(ob_ctx *c, int level, int b, double *M, double *dl, double *vol, int d,
                   double *fN, double *fA) {
  if (!c || level < 1 || level > c->nlev || b < 1 || b > c->nb) return OB_ERR_ARG;
  ob_blk *B = &c->dom[level - 1].blk[b - 1];
  if (M) memcpy(M, B->M, sizeof(double) * 9 * B->npad);
  if (dl) memcpy(dl, B->dl, sizeof(double) * 3 * B->npad);
  if (vol) memcpy(vol, B->vol, sizeof(double) * B->npad);
  if (d >= 1 && d <= 3) {
    if (fN) memcpy(fN, B->fN[d - 1], sizeof(double) * 3 * nfaces(B, d - 1));
    if (fA) memcpy(fA, B->fA[d - 1], sizeof(double) * nfaces(B, d - 1));
  }
  return 0;
}

int ob_wall(ob_ctx *c, int level, double *Tw, double *qw) {
  if (!c || level < 1 || level > c->nlev || !Tw || !qw) return OB_ERR_ARG;
  ob_dom *d = &c->dom[level - 1];
  /* ghosts mirrored from the current state: the reference reads the ghosts
   * its last stage left (mirrors of that stage's input), which only the 302
   * tangential differences at face corners see */
  ghost_fill(c, d);
  for (int n = 0; n < d->nbound; n++) wall_props(c, d, &d->bc[n], &Tw[n], &qw[n]);
  return c->err;
}

int ob_get_bc_ghost(ob_ctx *c, int level, int n, double *Mg1, double *matIDg, double *Tg) {
  if (!c || level < 1 || level > c->nlev) return OB_ERR_ARG;
  ob_dom *d = &c->dom[level - 1];
  if (n < 1 || n > d->nbound) return OB_ERR_ARG;
  ob_bc *e = &d->bc[n - 1];
  if (Mg1) memcpy(Mg1, e->Mg[0], 9 * sizeof(double));
  if (matIDg) *matIDg = e->matIDg;
  if (Tg) memcpy(Tg, e->Tg, 6 * sizeof(double));
  return 0;
}
