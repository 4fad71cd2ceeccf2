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

// BC_Connect_Metrics, Lib_Metrics.f90:704-936, first ghost layer of a 101
// entry: the ghost cell's eight nodes -- the four of the block's own face and
// four of the donor block's node layer beyond its connected face, at the
// donor stencil's offsets -- and their metric tensor Mg(1).  nm / ns: node
// (0,0,0) of the block / of the donor block ((n+1)^3 nodes, i fastest, three
// doubles each); q: the entry's 15 integers (b,i,j,k,f,type,bs,is,js,ks,fs,
// d11,d12,d21,d22).  Returns 0, 1 when a node lies outside its block (or the
// face is invalid), 2 when the metric is degenerate.
FUSS_HD int connect_metric_nodes(const double* nm, int nim, int njm, int nkm, const double* ns, int nis, int njs,
                                 int nks, const int* q, double* Mg1) {
  auto node = [](const double* base, int ni, int nj, int nk, int i, int j, int k) -> const double* {
    if (i < 0 || i > ni || j < 0 || j > nj || k < 0 || k > nk) return nullptr;
    return base + 3 * (i + (long)(ni + 1) * (j + (long)(nj + 1) * k));
  };
  auto Nm = [&](int i, int j, int k) { return node(nm, nim, njm, nkm, i, j, k); };
  auto Ns = [&](int i, int j, int k) { return node(ns, nis, njs, nks, i, j, k); };
  const int Im = q[1], Jm = q[2], Km = q[3], Fm = q[4];
  const int Is = q[7], Js = q[8], Ks = q[9], Fs = q[10], d11 = q[11], d12 = q[12], d21 = q[13], d22 = q[14];
  if (Fs < 1 || Fs > 6) return 1;
  const double* P0[8] = {nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr};
  const double* P1[8] = {nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr};
  switch (Fm) {
    case 1: P0[0] = Nm(Im - 1, Jm - 1, Km - 1); P0[1] = Nm(Im - 1, Jm - 1, Km); P0[2] = Nm(Im - 1, Jm, Km - 1); P0[3] = Nm(Im - 1, Jm, Km); break;
    case 2: P0[4] = Nm(Im, Jm - 1, Km - 1); P0[5] = Nm(Im, Jm - 1, Km); P0[6] = Nm(Im, Jm, Km - 1); P0[7] = Nm(Im, Jm, Km); break;
    case 3: P0[0] = Nm(Im - 1, Jm - 1, Km - 1); P0[1] = Nm(Im - 1, Jm - 1, Km); P0[4] = Nm(Im, Jm - 1, Km - 1); P0[5] = Nm(Im, Jm - 1, Km); break;
    case 4: P0[2] = Nm(Im - 1, Jm, Km - 1); P0[3] = Nm(Im - 1, Jm, Km); P0[6] = Nm(Im, Jm, Km - 1); P0[7] = Nm(Im, Jm, Km); break;
    case 5: P0[0] = Nm(Im - 1, Jm - 1, Km - 1); P0[2] = Nm(Im - 1, Jm, Km - 1); P0[4] = Nm(Im, Jm - 1, Km - 1); P0[6] = Nm(Im, Jm, Km - 1); break;
    case 6: P0[1] = Nm(Im - 1, Jm - 1, Km); P0[3] = Nm(Im - 1, Jm, Km); P0[5] = Nm(Im, Jm - 1, Km); P0[7] = Nm(Im, Jm, Km); break;
    default: return 1;
  }
  // g = 1: donor node index Is1 = Is + (gc-1) guides(Fs), the node layer on
  // the min side of a min face
  const int gsd = (Fs & 1) ? 0 : -1, gax = (Fs - 1) >> 1;
  const int Is1 = Is + (gax == 0 ? gsd : 0), Js1 = Js + (gax == 1 ? gsd : 0), Ks1 = Ks + (gax == 2 ? gsd : 0);
  int II, JJ, KK;
  if (Fs <= 2) { II = Is1; JJ = (2 * Js1 + d11 + d21) / 2; KK = (2 * Ks1 + d12 + d22) / 2; }
  else if (Fs <= 4) { II = (2 * Is1 + d11 + d21) / 2; JJ = Js1; KK = (2 * Ks1 + d12 + d22) / 2; }
  else { II = (2 * Is1 + d11 + d21) / 2; JJ = (2 * Js1 + d12 + d22) / 2; KK = Ks1; }
  // donor node shifted by -a1 along the first tangent and -a2 along the second
  auto donor = [&](int a1, int a2) -> const double* {
    if (Fs <= 2) return Ns(II, JJ - a1 * d11 - a2 * d21, KK - a1 * d12 - a2 * d22);
    if (Fs <= 4) return Ns(II - a1 * d11 - a2 * d21, JJ, KK - a1 * d12 - a2 * d22);
    return Ns(II - a1 * d11 - a2 * d21, JJ - a1 * d12 - a2 * d22, KK);
  };
  // slots of the four donor nodes per local face
  const int q1[6] = {4, 8, 6, 8, 7, 8}, q2[6] = {2, 6, 2, 4, 3, 4}, q3[6] = {3, 7, 5, 7, 5, 6}, q4[6] = {1, 5, 1, 3, 1, 2};
  P1[q1[Fm - 1] - 1] = donor(0, 0);
  P1[q2[Fm - 1] - 1] = donor(1, 0);
  P1[q3[Fm - 1] - 1] = donor(0, 1);
  P1[q4[Fm - 1] - 1] = donor(1, 1);
  // argument lists per face (Lib_Metrics.f90:192-229): g = 1 nodes in P1, g - 1 = 0 in P0
  const double* n1[8];
  switch (Fm) {
    case 1: n1[0] = P1[0]; n1[1] = P1[1]; n1[2] = P1[2]; n1[3] = P1[3]; n1[4] = P0[0]; n1[5] = P0[1]; n1[6] = P0[2]; n1[7] = P0[3]; break;
    case 2: n1[0] = P0[4]; n1[1] = P0[5]; n1[2] = P0[6]; n1[3] = P0[7]; n1[4] = P1[4]; n1[5] = P1[5]; n1[6] = P1[6]; n1[7] = P1[7]; break;
    case 3: n1[0] = P1[0]; n1[1] = P1[1]; n1[2] = P0[0]; n1[3] = P0[1]; n1[4] = P1[4]; n1[5] = P1[5]; n1[6] = P0[4]; n1[7] = P0[5]; break;
    case 4: n1[0] = P0[2]; n1[1] = P0[3]; n1[2] = P1[2]; n1[3] = P1[3]; n1[4] = P0[6]; n1[5] = P0[7]; n1[6] = P1[6]; n1[7] = P1[7]; break;
    case 5: n1[0] = P1[0]; n1[1] = P0[0]; n1[2] = P1[2]; n1[3] = P0[2]; n1[4] = P1[4]; n1[5] = P0[4]; n1[6] = P1[6]; n1[7] = P0[6]; break;
    default: n1[0] = P0[1]; n1[1] = P1[1]; n1[2] = P0[3]; n1[3] = P1[3]; n1[4] = P0[5]; n1[5] = P1[5]; n1[6] = P0[7]; n1[7] = P1[7]; break;
  }
  for (int s = 0; s < 8; s++)
    if (!n1[s]) return 1;
  double dl[3], v;
  return metric_tensor(n1[0], n1[1], n1[2], n1[3], n1[4], n1[5], n1[6], n1[7], Mg1, dl, &v) ? 2 : 0;
}

}  // namespace fuss
