// Shared device helpers for the FastPM particle-mesh kernels (gfx950 / CDNA4).
//
// Conventions used by every kernel in this library:
//  * Real meshes are C-ordered [n0][n1][n2] with no padding.
//  * Half-spectrum complex meshes are [n0][n1][pitch2] with pitch2 >= n2/2+1;
//    mode (i, j, k) holds k_x = 2*pi*fftfreq(i), k_y = 2*pi*fftfreq(j),
//    k_z = 2*pi*k over the box side.  The forward transform carries 1/(n0 n1 n2).
//  * Particle positions are split as X = Q + S where Q is either an implicit
//    Lagrangian lattice (particle p -> lattice cell (p/(ny nz), (p/nz)%ny, p%nz))
//    or an explicit [n][3] array, and S is the displacement [n][3].
//  * A grid point (i, j, k) sits at position (i, j, k) * BoxSize / Nmesh.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>
#include "../../include/fpm.h"

namespace fpm {

template <typename T> struct cvec;
template <> struct cvec<float> { using type = float2; };
template <> struct cvec<double> { using type = double2; };
template <typename T> using C = typename cvec<T>::type;

template <typename V> __host__ __device__ __forceinline__ V cmk(decltype(V::x) a, decltype(V::x) b) { V r; r.x = a; r.y = b; return r; }
template <typename V> __device__ __forceinline__ V cadd(V a, V b) { return cmk<V>(a.x + b.x, a.y + b.y); }
template <typename V> __device__ __forceinline__ V csub(V a, V b) { return cmk<V>(a.x - b.x, a.y - b.y); }
template <typename V> __device__ __forceinline__ V cmul(V a, V b) { return cmk<V>(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x); }
// Complex product with its rounding fixed: each component is one explicit
// fused multiply-add onto a rounded product, which the compiler cannot
// contract in another way, so two kernels inlining the same body (a box's own
// launch and a batched one) round it alike.
__device__ __forceinline__ float fmadd(float a, float b, float c) { return __builtin_fmaf(a, b, c); }
__device__ __forceinline__ double fmadd(double a, double b, double c) { return __builtin_fma(a, b, c); }
template <typename V> __device__ __forceinline__ V cmul_fma(V a, V b) {
  return cmk<V>(fmadd(a.x, b.x, -(a.y * b.y)), fmadd(a.x, b.y, a.y * b.x));
}
template <typename V> __device__ __forceinline__ V cconj(V a) { return cmk<V>(a.x, -a.y); }
template <typename V> __device__ __forceinline__ V cscale(V a, decltype(V::x) s) { return cmk<V>(a.x * s, a.y * s); }

// Map a linear block id so that consecutive tiles land on the same XCD.
// Blocks are dispatched round-robin over the 8 XCDs (block b -> XCD b % 8);
// this bijection gives XCD x a contiguous range of tiles, so neighbouring tiles
// share that XCD's L2.
__device__ __forceinline__ int64_t xcd_remap(int64_t bid, int64_t nblk) {
  const int64_t q = nblk >> 3, r = nblk & 7;
  const int64_t x = bid & 7, k = bid >> 3;
  return x * q + (x < r ? x : r) + k;
}

// Linear index of mesh cell (x, y, z), each coordinate within its axis: the
// row x * n1 + y by a full-rate 24-bit multiply-add (both factors are below
// 2^24 and the row count n0 * n1 below 2^32), then one 32 x 32 -> 64-bit
// multiply-add for the row's first cell.  The general 64-bit form costs
// three quarter-rate multiplies.
__device__ __forceinline__ int64_t mesh_index(int x, int y, int z, const int nm[3]) {
  const uint32_t row = __umul24((uint32_t)x, (uint32_t)nm[1]) + (uint32_t)y;
  return (int64_t)((uint64_t)row * (uint32_t)nm[2] + (uint32_t)z);
}

// Periodic wrap of a cell index into [0, n).  Indices within one period of
// the box (every particle the solver moves) take the compare-and-add paths;
// only others pay for the 64-bit remainder.
__device__ __forceinline__ int64_t wrapi(int64_t i, int64_t n) {
  if (i >= 0 && i < n) return i;
  if (i < 0 && i >= -n) return i + n;
  if (i >= n && i < 2 * n) return i - n;
  const int64_t r = i % n;
  return r < 0 ? r + n : r;
}

// Particle position in mesh-cell units along one axis, returned as an exact
// integer cell plus a fractional offset in [0, 1).  For the implicit lattice the
// integer part comes from integer arithmetic, so precision is set by the
// displacement alone, not by the box size.
template <typename T>
struct PartDesc {
  const T* q;         // explicit Lagrangian positions [n][3] or nullptr
  const T* s;         // displacement [n][3] (may be nullptr: zero)
  int64_t n;
  int32_t lat[3];     // lattice particles per side
  double lat_shift;   // lattice offset in lattice spacings
  double ratio[3];    // mesh cells per lattice spacing (Nmesh / Nlattice)
  T inv_cell[3];      // Nmesh / BoxSize
};

// A particle's three components, moved as one 12-byte access.
template <typename T>
struct Vec3 {
  T v[3];
};

// Lattice coordinates of lattice particle p (C order over P.lat).
template <typename T>
__device__ __forceinline__ void lattice_of(const PartDesc<T>& P, int64_t p, int32_t li[3]) {
  if (p <= 0x7fffffffLL) {  // 32-bit index arithmetic (far cheaper than 64-bit division)
    const uint32_t nz = (uint32_t)P.lat[2], ny = (uint32_t)P.lat[1];
    const uint32_t pu = (uint32_t)p;
    const uint32_t t = pu / nz;
    li[2] = (int32_t)(pu - t * nz);
    li[0] = (int32_t)(t / ny);
    li[1] = (int32_t)(t - (uint32_t)li[0] * ny);
  } else {
    const int64_t nz = P.lat[2], ny = P.lat[1];
    const int64_t t = p / nz;
    li[2] = (int32_t)(p - t * nz);
    li[0] = (int32_t)(t / ny);
    li[1] = (int32_t)(t - (int64_t)li[0] * ny);
  }
}

// Cell and fraction of a lattice particle at lattice coordinates li whose
// displacement is sv (in registers).  No FMA contraction: every kernel that
// places a particle (binning, paint, readout) must round its position
// identically.
template <typename T>
__device__ __forceinline__ void lattice_cell_s(const PartDesc<T>& P, const int32_t li[3],
                                               const T sv[3], int64_t cell[3], T frac[3]) {
#pragma clang fp contract(off)
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    const double base = ((double)li[d] + P.lat_shift) * P.ratio[d];
    const double fb = floor(base);
    T t2 = (T)(base - fb);
    t2 += sv[d] * P.inv_cell[d];
    const T ft = floor(t2);
    cell[d] = (int64_t)(int32_t)fb + (int64_t)(int32_t)ft;
    frac[d] = t2 - ft;
  }
}

// The same for lattice particle p, its displacement read from P.s (or zero).
template <typename T>
__device__ __forceinline__ void lattice_cell(const PartDesc<T>& P, int64_t p, const int32_t li[3],
                                             int64_t cell[3], T frac[3]) {
#pragma clang fp contract(off)
  if (P.s) {  // the displacement as one 12-byte load
    const Vec3<T> sv = reinterpret_cast<const Vec3<T>*>(P.s)[p];
    lattice_cell_s(P, li, sv.v, cell, frac);
    return;
  }
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    const double base = ((double)li[d] + P.lat_shift) * P.ratio[d];
    const double fb = floor(base);
    const T t2 = (T)(base - fb);
    const T ft = floor(t2);
    cell[d] = (int64_t)(int32_t)fb + (int64_t)(int32_t)ft;
    frac[d] = t2 - ft;
  }
}

// Cells are exact for positions within 2^31 cells of the box.
template <typename T>
__device__ __forceinline__ void particle_cell(const PartDesc<T>& P, int64_t p, int64_t cell[3], T frac[3]) {
#pragma clang fp contract(off)
  if (P.q == nullptr) {
    int32_t li[3];
    lattice_of(P, p, li);
    lattice_cell(P, p, li, cell, frac);
  } else {
#pragma unroll
    for (int d = 0; d < 3; ++d) {
      T x = P.q[3 * p + d];
      if (P.s) x += P.s[3 * p + d];
      x *= P.inv_cell[d];
      const T fx = floor(x);
      cell[d] = (int64_t)(int32_t)fx;
      frac[d] = x - fx;
    }
  }
}

// y + a x rounded once: every kick and drift update (alone, table-driven or
// fused) uses this one expression, so fused and separate updates agree bit
// for bit.
__device__ __forceinline__ float madd(float a, float x, float y) { return __builtin_fmaf(a, x, y); }
__device__ __forceinline__ double madd(double a, double x, double y) { return __builtin_fma(a, x, y); }

// particle_cell with the displacement given (sv[d]) instead of read from P.s.
template <typename T>
__device__ __forceinline__ void particle_cell_s(const PartDesc<T>& P, int64_t p, const T sv[3],
                                                int64_t cell[3], T frac[3]) {
#pragma clang fp contract(off)
  if (P.q == nullptr) {
    int32_t li[3];
    lattice_of(P, p, li);
    lattice_cell_s(P, li, sv, cell, frac);
  } else {
#pragma unroll
    for (int d = 0; d < 3; ++d) {
      T x = P.q[3 * p + d];
      x += sv[d];
      x *= P.inv_cell[d];
      const T fx = floor(x);
      cell[d] = (int64_t)(int32_t)fx;
      frac[d] = x - fx;
    }
  }
}

// A step boundary's kicks and drift for one particle, fused into a pass that
// then needs its position: P += k0 F [+ k1 F] (nkick of them), S += d P, each
// element rounded as fpm_kick_drift rounds it.  S == nullptr: none.
template <typename T>
struct KickDrift {
  T* S;
  T* P;
  const T* F;
  T k0, k1, d;
  int nkick;
};

template <typename T>
__device__ __forceinline__ void kick_drift_one(const KickDrift<T>& kd, int64_t p, T sv[3]) {
  Vec3<T>* P3 = reinterpret_cast<Vec3<T>*>(kd.P);
  Vec3<T>* S3 = reinterpret_cast<Vec3<T>*>(kd.S);
  Vec3<T> pv = P3[p];
  const Vec3<T> s0 = S3[p];
  if (kd.nkick > 0) {
    const Vec3<T> f = reinterpret_cast<const Vec3<T>*>(kd.F)[p];
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      pv.v[c] = madd(kd.k0, f.v[c], pv.v[c]);
      if (kd.nkick > 1) pv.v[c] = madd(kd.k1, f.v[c], pv.v[c]);
    }
    P3[p] = pv;
  }
  Vec3<T> s1;
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    s1.v[c] = madd(kd.d, pv.v[c], s0.v[c]);
    sv[c] = s1.v[c];
  }
  S3[p] = s1;
}

}  // namespace fpm

#define FPM_LAUNCH_CHECK()                                   \
  do {                                                       \
    hipError_t e_ = hipGetLastError();                       \
    if (e_ != hipSuccess) return FPM_ERR_HIP;                \
  } while (0)
