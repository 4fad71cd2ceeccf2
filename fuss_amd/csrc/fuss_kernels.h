// fuss_kernels.h -- device data structures and kernels of the FUSS explicit
// step on gfx950 (MI355X).  Included by fuss_gpu.hip only.
//
// Layout (per multigrid level, all blocks of the level in one arena):
//   cell arrays are padded by GC=2 ghost layers in i and j and, in 3-D, in k
//   (2-D levels keep a single k plane on each side, which only stores the
//   k-face geometry: the reference zeroes every k term when ndir == 2,
//   Lib_Diffusive.f90:32).  Cell (i,j,k) of block b lives at
//   blk[b].base + (i-1) + (j-1)*sy + (k-1)*sz (doubles), i fastest; sy is
//   ni+4 rounded up to ROW_ALIGN and base is a multiple of ROW_ALIGN, so
//   every interior row starts on a 128-byte line.
//   Face weights: W[3*D+c][cell] holds, for the face between cell and its
//   +1 neighbour in direction D, A * sum_x Mbar(c,x) n_x with
//   Mbar = (M_cell + M_nb)/2 -- the flux is then kappa * sum_c G_c W[3D+c].
#pragma once

#include "../../include/fuss_gpu.h"
#include "fuss_common.h"

namespace fuss {

constexpr int TILE_X = 64;
constexpr int TILE_Y = 4;
constexpr int TILE_T = TILE_X * TILE_Y;
constexpr int ROW_ALIGN = 16;        // doubles per 128-byte line
constexpr int MAX_SERIES = 16;
constexpr int RUN_G = 30;            // steps per captured graph of fuss_run (multiple of 3)

struct BlockInfo {
  int ni, nj, nk;
  int sy;            // ni + 2*GC
  long sz;           // sy * (nj + 2*GC)
  long off;          // arena index of the padded origin
  long base;         // arena index of cell (1,1,1)
  long npad;         // padded size
  long node_off;     // node arena offset (in nodes)
  int fmap_off[6];   // face-map offsets, -1 when the face has no entries
  int face_active;   // bit f-1 set when face f has flux-producing entries
  int tang;          // nonzero when any tangential weight is nonzero
  int umat;          // material index when the whole block has one material, else -1
  int conn;          // bit f-1 set when face f has a connection (101) entry
  int uqv;           // nonzero when every cell of the block has the same qvol
  int roff;          // rectilinear block: offset of its per-axis metric factors, else -1
};

struct Tile {
  int b, i0, j0, k0;
};

// Per-axis metric factors of a rectilinear block (every node lies on the
// coordinate planes x = x_i, y = y_j, z = z_k).  Nine arrays of n_d + 2
// values start at BlockInfo::roff: (a, f, s) for d = x, y, z, indexed by the
// cell index q = 0 .. n_d + 1 (0 and n_d + 1 repeat the end cells):
//   a_d(q)  edge length of the cells of index q along d (node difference
//           x_q - x_{q-1}, as the face areas and tetrahedra use it),
//   f_d(q)  normal coefficient of the face between q and q+1,
//           0.5 (M_dd(q) + M_dd(q+1)) N_d (0 on the block's faces),
//   s_d(q)  1 / dl_d^2 (dl_d of Compute_Metric_Tensor depends on q alone).
// Cell volume (a_x a_y) a_z, normal flux weights (f_x a_y) a_z,
// (f_y a_x) a_z and f_z (a_x a_y), sum of 1/dl^2 (s_x + s_y) + s_z.
struct RectAxes {
  const double *ax, *fx, *qx, *ay, *fy, *qy, *az, *fz, *qz;  // q: 1/dl^2
};

__host__ __device__ inline RectAxes rect_axes(const double* rect, const BlockInfo& B) {
  const long nx = B.ni + 2, ny = B.nj + 2, nz = B.nk + 2;
  const double* p = rect + B.roff;
  RectAxes r;
  r.ax = p; r.fx = p + nx; r.qx = p + 2 * nx;
  p += 3 * nx;
  r.ay = p; r.fy = p + ny; r.qy = p + 2 * ny;
  p += 3 * ny;
  r.az = p; r.fz = p + nz; r.qz = p + 2 * nz;
  return r;
}

// Time series of 301/302 values on the device (one concatenated table).
struct SerDev {
  const double* t;
  const double* v;
  const int* off;      // first point of series s
  const int* n;        // points of series s
  const int* per;      // periodic flag
};

// One BC entry, geometry gathered at setup.  Reference fields: bc_type /
// FUSS_bc_type (src/lib/base/Advanced_Types_m.f90:83-114).
struct BCRec {
  int type, f, dir, series;
  int b, i, j, k;      // 0-based block, 1-based cell
  long cell;           // arena index of the boundary cell
  long gdst, gsrc;     // layer-1 ghost of the cell and its source (mirror or donor)
  long dn[4];          // donor cells of Tg(3..6) (connections)
  long rdonor;         // donor cell whose R fills the IRS ghost (connections)
  int matg;           // material of the connection ghost (matIDg)
  int fpos;           // position of the entry's face cell in the level's face map
  double qw, Tw, hconv, Tref, eps;
  double A, N[3];
  double M[9];         // metric tensor of the boundary cell, c(3,3) column-major
  double Mg[9];        // ghost-cell metric Mg(1) (connections)
  double mc[3];        // dot(Mg(r,:), N), r = 1..3
};

// The per-stage fields of a wall entry of type 301, 303, 304 or 305 in one
// 80-byte record (the full BCRec is ~380 bytes): the surface flux pass reads
// these instead of whole records.  Filled on the device from the BCRec
// array after its geometry (k_bc_hot) and again after fuss_set_bc_values.
struct BCHot {
  int type, f, m, series;  // BC type, face, material of the cell, time series (-1: none)
  unsigned cell;           // arena index of the boundary cell
  int b;
  int fpos, pad_;          // face-map position of the entry
  double A;                // face area
  double mn;               // dot(M(dir,:), N) of the boundary cell
  double hconv, Tref, qw, eps;
};

// Property tables, one row per kelvin from Ti1 (Load_Table.f90:61-71).
template <typename R>
struct Tab {
  const R* cp;
  const R* rho;
  const R* kap;
  const R* h;
  // the same tables as (v(T), v(T+1) - v(T)) rows: one vector load per lookup
  const R* kap2;       // 2 per row
  const R* pr4;        // rho, drho, cp, dcp per row
  int nT, Ti1, Ti2;
};

template <typename R> struct Vec2;
template <> struct Vec2<double> { using T = double2; };
template <> struct Vec2<float> { using T = float2; };
template <typename R> struct Vec4;
template <> struct Vec4<double> { using T = double4; };
template <> struct Vec4<float> { using T = float4; };

// 32-bit element offsets from a uniform base: lets the compiler use the
// SGPR-base + 32-bit VGPR-offset form of global loads and stores.
template <typename T>
__device__ __forceinline__ T ldu(const T* __restrict__ base, unsigned idx) {
  return *reinterpret_cast<const T*>(reinterpret_cast<const char*>(base) + idx * (unsigned)sizeof(T));
}
// (a byte offset instead of an element index)
template <typename T>
__device__ __forceinline__ T ldb(const void* base, unsigned off) {
  return *reinterpret_cast<const T*>(reinterpret_cast<const char*>(base) + off);
}
template <typename T>
__device__ __forceinline__ void stu(T* __restrict__ base, unsigned idx, T v) {
  *reinterpret_cast<T*>(reinterpret_cast<char*>(base) + idx * (unsigned)sizeof(T)) = v;
}
// Streamed (read-once / write-once) elements: non-temporal accesses, so that
// they do not evict the property-table rows every lookup reuses from the
// caches.
template <typename T>
__device__ __forceinline__ T ldn(const T* __restrict__ base, unsigned idx) {
  return __builtin_nontemporal_load(
      reinterpret_cast<const T*>(reinterpret_cast<const char*>(base) + idx * (unsigned)sizeof(T)));
}
template <typename T>
__device__ __forceinline__ void stn(T* __restrict__ base, unsigned idx, T v) {
  __builtin_nontemporal_store(v, reinterpret_cast<T*>(reinterpret_cast<char*>(base) + idx * (unsigned)sizeof(T)));
}

// Interpolation from a (v, dv) row: v(Tm) + (T - Tm) * (v(Tm+1) - v(Tm)),
// the same operations as tab_interp (the difference is formed once at setup).
template <typename R>
__device__ __forceinline__ R tab_interp2(const R* __restrict__ t2, R T, int Ti1, int Ti2, int* bad) {
  if (!(T >= (R)Ti1 && T < (R)Ti2)) {
    *bad = 1;
    T = T < (R)Ti1 || T != T ? (R)Ti1 : (R)(Ti2 - 1);
  }
  const int Tm = (int)T;
  const typename Vec2<R>::T v = ldu(reinterpret_cast<const typename Vec2<R>::T*>(t2), (unsigned)(Tm - Ti1));
  return v.x + (T - (R)Tm) * v.y;
}

__device__ __forceinline__ void flag_error(int* st, int code, int b, int i, int j, int k) {
  if (atomicCAS(st, 0, code) == 0) {
    st[1] = b + 1;
    st[2] = i;
    st[3] = j;
    st[4] = k;
  }
}

// Check_And_Fix_State failure (Lib_Newstate.f90:278-290): the reference
// stops at the first failing cell of its loops (stage, then block, k, j, i).
// Every failing cell offers the packed key (stage, b, k, j, i) to an atomic
// minimum at status[6..7] (initialised to all ones), so the reported cell is
// that first one whatever order the work-groups run in.
__device__ __forceinline__ void flag_state(int* st, int irk, int b, int i, int j, int k) {
  atomicCAS(st, 0, FUSS_ESTATE);
  const unsigned long long key = ((unsigned long long)(irk & 7) << 61) | ((unsigned long long)((b + 1) & 0x1fff) << 48) |
                                 ((unsigned long long)(k & 0xffff) << 32) | ((unsigned long long)(j & 0xffff) << 16) |
                                 (unsigned long long)(i & 0xffff);
  atomicMin(reinterpret_cast<unsigned long long*>(st + 6), key);
}

// Workgroups are dealt to the 8 XCDs round-robin; give each XCD a
// contiguous run of tiles so that neighbouring k-planes share its L2.
__device__ __forceinline__ int xcd_remap(int bid, int n) {
  const int q = n >> 3, r = n & 7;
  const int x = bid & 7, s = bid >> 3;
  return x * q + (x < r ? x : r) + s;
}

// co_Kappa / co_Rho / co_Cp / co_H, Lib_Solid.f90:11-76:
// v = v(Tm) + (T - Tm) * (v(Tm+1) - v(Tm)), Tm = int(T).
template <typename R>
__device__ __forceinline__ R tab_interp(const R* __restrict__ row, R T, int Ti1, int Ti2,
                                        int* bad) {
  if (!(T >= (R)Ti1 && T < (R)Ti2)) {
    *bad = 1;
    T = T < (R)Ti1 || T != T ? (R)Ti1 : (R)(Ti2 - 1);
  }
  const int Tm = (int)T;
  const R vm = row[Tm - Ti1], vp = row[Tm - Ti1 + 1];
  return vm + (T - (R)Tm) * (vp - vm);
}
// co_DrhoDT / co_DcpDT, Lib_Solid.f90:101-132
template <typename R>
__device__ __forceinline__ R tab_slope(const R* __restrict__ row, R T, int Ti1, int Ti2) {
  if (!(T >= (R)Ti1 && T < (R)Ti2)) T = T < (R)Ti1 || T != T ? (R)Ti1 : (R)(Ti2 - 1);
  const int Tm = (int)T;
  return row[Tm - Ti1 + 1] - row[Tm - Ti1];
}
// co_T, Lib_Solid.f90:79-98: the first row s with h(s) > hv brackets
// [s-1, s].  The reference scans from the first row; on a non-decreasing
// table the local walk from the previous temperature finds the same row.
template <typename R>
__device__ __forceinline__ R tab_inverse(const R* __restrict__ row, R hv, R Tguess, int Ti1,
                                         int Ti2, int* bad) {
  int s = (Tguess >= (R)Ti1 && Tguess < (R)Ti2) ? (int)Tguess + 1 : Ti1 + 1;
  // invariant sought: h(s-1) <= hv < h(s), Ti1+1 <= s <= Ti2
  int guard = 0;
  while (s > Ti1 + 1 && row[s - 1 - Ti1] > hv && guard < 1 << 20) {
    --s;
    ++guard;
  }
  while (s <= Ti2 && !(row[s - Ti1] > hv) && guard < 1 << 20) {
    ++s;
    ++guard;
  }
  if (s > Ti2 || !(row[s - 1 - Ti1] <= hv) || hv != hv) {
    *bad = 1;
    s = s > Ti2 ? Ti2 : (s < Ti1 + 1 ? Ti1 + 1 : s);
  }
  const R hm = row[s - 1 - Ti1], hp = row[s - Ti1];
  return (R)(s - 1) + (hv - hm) / (hp - hm);
}

// co_T by bisection over the whole table (same first row s with h(s) > hv).
template <typename R>
__device__ __forceinline__ R tab_inverse_search(const R* __restrict__ row, R hv, int Ti1, int Ti2, int* bad) {
  int lo = Ti1 + 1, hi = Ti2;
  if (!(row[Ti2 - Ti1] > hv) || !(row[0] <= hv)) {
    *bad = 1;
    return (R)Ti1;
  }
  while (lo < hi) {
    const int mid = (lo + hi) >> 1;
    if (row[mid - Ti1] > hv) hi = mid;
    else lo = mid + 1;
  }
  const R hm = row[lo - 1 - Ti1], hp = row[lo - Ti1];
  return (R)(lo - 1) + (hv - hm) / (hp - hm);
}

// time_series_type%update, src/lib/base/Series_Data_m.f90:55-86
__host__ __device__ inline double series_value(int periodic, int n, const double* t, const double* v, double time) {
  double t_ = time;
  if (periodic) {
    const double period = t[n - 1] - t[0];
    if (period > 0.0) {
      const double a = t_ - t[0];
      t_ = t[0] + (a - floor(a / period) * period);
    }
  }
  if (t_ == t[0]) return v[0];
  for (int i = 1; i < n; i++)
    if (t_ > t[i - 1] && t_ <= t[i]) return (v[i] - v[i - 1]) / (t[i] - t[i - 1]) * (t_ - t[i - 1]) + v[i - 1];
  return t_ > t[n - 1] ? v[n - 1] : 0.0;
}

}  // namespace fuss
