// fuss_common.h -- host/device helpers shared by the FUSS gfx950 kernels.
//
// Geometry routines restate src/lib/numerics/space/Lib_Metrics.f90 in the
// reference's operation order so that the device-computed metrics equal the
// reference's bit for bit (the file is compiled with FP contraction off for
// these functions).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#define FUSS_HD __host__ __device__ __forceinline__

namespace fuss {

constexpr int GC = 2;  // ghost layers, src/lib/base/Global_m.f90:9
constexpr double SIGMA_SB = 5.67e-8;  // src/lib/base/Parameters_m.f90:10

// guide(6,3) face offsets, src/lib/base/Parameters_m.f90:11-16
FUSS_HD int guide(int f, int d) {
  // f = 1..6, d = 0..2
  const int axis = (f - 1) >> 1;
  const int s = ((f - 1) & 1) ? -1 : 1;
  return axis == d ? s : 0;
}

// Compute_Metric_Tensor, Lib_Metrics.f90:250-376.  Nodes are 3-vectors.
// M is written column-major c(3,3): c(r,c) at M[(r-1)+3*(c-1)].
// Returns 0, or 1 for det=0, 2 for vol<=0.
FUSS_HD double tet_vol(const double* vx, const double* vy, const double* vz, int i1, int i2,
                       int i3, int i4) {
#pragma clang fp contract(off)
  i1--; i2--; i3--; i4--;
  double v = ((vx[i2] - vx[i1]) *
                  ((vy[i3] - vy[i1]) * (vz[i4] - vz[i1]) - (vy[i4] - vy[i1]) * (vz[i3] - vz[i1])) +
              (vy[i2] - vy[i1]) *
                  ((vx[i4] - vx[i1]) * (vz[i3] - vz[i1]) - (vx[i3] - vx[i1]) * (vz[i4] - vz[i1])) +
              (vz[i2] - vz[i1]) *
                  ((vx[i3] - vx[i1]) * (vy[i4] - vy[i1]) - (vx[i4] - vx[i1]) * (vy[i3] - vy[i1]))) /
             6.0;
  return v < 0.0 ? -v : v;
}

FUSS_HD int metric_tensor(const double* N1, const double* N2, const double* N3, const double* N4,
                          const double* N5, const double* N6, const double* N7, const double* N8,
                          double* M, double* dl, double* vol) {
#pragma clang fp contract(off)
  double A[3][3];
  for (int x = 0; x < 3; x++) {
    A[0][x] = N5[x] - N1[x] + N6[x] - N2[x] + N7[x] - N3[x] + N8[x] - N4[x];
    A[1][x] = N3[x] - N1[x] + N4[x] - N2[x] + N7[x] - N5[x] + N8[x] - N6[x];
    A[2][x] = N2[x] - N1[x] + N4[x] - N3[x] + N6[x] - N5[x] + N8[x] - N7[x];
  }
  for (int r = 0; r < 3; r++)
    for (int x = 0; x < 3; x++) A[r][x] = 0.25 * A[r][x];
  const double det = A[0][0] * A[1][1] * A[2][2] - A[0][0] * A[1][2] * A[2][1] -
                     A[0][1] * A[1][0] * A[2][2] + A[0][1] * A[1][2] * A[2][0] +
                     A[0][2] * A[1][0] * A[2][1] - A[0][2] * A[1][1] * A[2][0];
  int err = 0;
  if (det == 0.0) {
    err = 1;
    for (int q = 0; q < 9; q++) M[q] = (q % 4 == 0) ? 1.0 : 0.0;
  } else {
    double cof[3][3];
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
      for (int x = 0; x < 3; x++) M[r + 3 * x] = cof[r][x] / det;
  }
  for (int h = 0; h < 3; h++) dl[h] = sqrt(A[h][0] * A[h][0] + A[h][1] * A[h][1] + A[h][2] * A[h][2]);
  // vertex order of the 5-tetrahedra split (Lib_Metrics.f90:317-351)
  const double* P[8] = {N1, N5, N3, N7, N2, N6, N4, N8};
  double vx[8], vy[8], vz[8];
  for (int q = 0; q < 8; q++) {
    vx[q] = P[q][0];
    vy[q] = P[q][1];
    vz[q] = P[q][2];
  }
  *vol = tet_vol(vx, vy, vz, 1, 2, 3, 5) + tet_vol(vx, vy, vz, 2, 4, 3, 8) +
         tet_vol(vx, vy, vz, 5, 8, 6, 2) + tet_vol(vx, vy, vz, 5, 7, 8, 3) +
         tet_vol(vx, vy, vz, 5, 8, 2, 3);
  if (!err && !(*vol > 0.0)) err = 2;
  return err;
}

// Face area vector of the quadrilateral from its two diagonals d1 x d2 / 2
// (Compute_Norm_Area, Lib_Metrics.f90:159-244).
FUSS_HD void face_normal(const double* p, const double* q, const double* r, const double* s,
                         double sign, double* N, double* A) {
#pragma clang fp contract(off)
  double d1[3], d2[3], d3[3];
  for (int x = 0; x < 3; x++) {
    d1[x] = p[x] - q[x];
    d2[x] = r[x] - s[x];
  }
  d3[0] = (d1[1] * d2[2] - d1[2] * d2[1]) * .5;
  d3[1] = (d1[2] * d2[0] - d1[0] * d2[2]) * .5;
  d3[2] = (d1[0] * d2[1] - d1[1] * d2[0]) * .5;
  const double a = sqrt(d3[0] * d3[0] + d3[1] * d3[1] + d3[2] * d3[2]);
  N[0] = d3[0] / a * sign;
  N[1] = d3[1] / a * sign;
  N[2] = d3[2] / a * sign;
  if (a == 0.0) N[0] = N[1] = N[2] = 0.0;
  *A = a;
}

}  // namespace fuss
