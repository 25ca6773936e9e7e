// Kernels for the reverse-mode (adjoint) pass of the PM solver.
//
//   paint_weighted : mass assignment with a per-particle mass (the adjoint of
//                    readout with respect to the mesh).
//   window_grad    : sum_c w_c[p] * sum_cells mesh_c(cell) dW/dx (the position
//                    derivative of readout and of paint).
//   kspace_accum   : acc (+)= M(k) * in or conj(M(k)) * in for one k-space
//                    multiplier M (the adjoint of a multiplier is its conjugate).
//   hermitian_scale: scales the modes with 0 < k_z < n2/2 (stored once, counted
//                    twice in a real field) and the k_z = 0, n2/2 planes.
//   lpt2_source_vjp: pointwise derivative of the 2LPT source with respect to
//                    the six second derivatives.
#include "fpm_common.h"

namespace fpm {

template <typename T>
int make_desc(const fpm_particles* part, const fpm_mesh* m, PartDesc<T>& P);

static unsigned grid_n(int64_t n, int64_t cap = 1 << 20) {
  int64_t g = (n + 255) / 256;
  if (g < 1) g = 1;
  if (g > cap) g = cap;
  return (unsigned)g;
}

// threads of a per-row workgroup: the row length rounded up to a whole
// wave, at most 256
static unsigned row_threads(int n2) {
  const int t = (n2 + 63) / 64 * 64;
  return (unsigned)(t > 256 ? 256 : (t < 64 ? 64 : t));
}

// Window origin, weights and weight derivatives (per unit of the mesh-cell
// coordinate) for one axis.
template <typename T, int ORDER>
__device__ __forceinline__ void window_d(int64_t cell, T f, int n, int64_t& i0, T w[ORDER],
                                         T dw[ORDER]) {
  int64_t o;
  if constexpr (ORDER == 2) {
    o = cell;
    w[0] = (T)1 - f;
    w[1] = f;
    dw[0] = (T)-1;
    dw[1] = (T)1;
  } else {
    const bool up = f >= (T)0.5;
    const T d = up ? f - (T)1 : f;
    o = cell + (up ? 1 : 0) - 1;
    const T a = (T)0.5 - d, b = (T)0.5 + d;
    w[0] = (T)0.5 * a * a;
    w[1] = (T)0.75 - d * d;
    w[2] = (T)0.5 * b * b;
    dw[0] = -a;
    dw[1] = (T)-2 * d;
    dw[2] = b;
  }
  i0 = wrapi(o, n);
}

template <typename T, int ORDER>
__global__ __launch_bounds__(256) void k_paint_weighted(PartDesc<T> P, T* mesh, int n0, int n1,
                                                         int n2, const T* wts, int64_t wstride,
                                                         T scale) {
  const int nm[3] = {n0, n1, n2};
  for (int64_t p = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; p < P.n;
       p += (int64_t)gridDim.x * blockDim.x) {
    int64_t cell[3];
    T f[3];
    particle_cell(P, p, cell, f);
    int64_t i0[3];
    T w[3][ORDER], dw[3][ORDER];
#pragma unroll
    for (int d = 0; d < 3; ++d) window_d<T, ORDER>(cell[d], f[d], nm[d], i0[d], w[d], dw[d]);
    const T mass = scale * wts[p * wstride];
#pragma unroll
    for (int a = 0; a < ORDER; ++a) {
      int64_t ix = i0[0] + a;
      if (ix >= n0) ix -= n0;
#pragma unroll
      for (int b = 0; b < ORDER; ++b) {
        int64_t iy = i0[1] + b;
        if (iy >= n1) iy -= n1;
        const T wab = w[0][a] * w[1][b] * mass;
        T* row = mesh + (ix * n1 + iy) * n2;
#pragma unroll
        for (int c = 0; c < ORDER; ++c) {
          int64_t iz = i0[2] + c;
          if (iz >= n2) iz -= n2;
          atomicAdd(row + iz, wab * w[2][c]);
        }
      }
    }
  }
}

template <typename T>
struct GradArgs {
  const T* mesh[4];
  const T* wts[4];     // per-particle weights, or nullptr for 1
  int64_t wstride[4];
  int ncomp;
  T scale;
  T* gx;               // [n][3]
  int accumulate;
};

template <typename T, int ORDER>
__global__ __launch_bounds__(256) void k_window_grad(PartDesc<T> P, GradArgs<T> G, int n0, int n1,
                                                      int n2) {
  const int nm[3] = {n0, n1, n2};
  for (int64_t p = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; p < P.n;
       p += (int64_t)gridDim.x * blockDim.x) {
    int64_t cell[3];
    T f[3];
    particle_cell(P, p, cell, f);
    int64_t i0[3];
    T w[3][ORDER], dw[3][ORDER];
#pragma unroll
    for (int d = 0; d < 3; ++d) window_d<T, ORDER>(cell[d], f[d], nm[d], i0[d], w[d], dw[d]);
    T g[3] = {(T)0, (T)0, (T)0};
    for (int c = 0; c < G.ncomp; ++c) {
      const T wt = G.wts[c] ? G.wts[c][p * G.wstride[c]] : (T)1;
      const T* m = G.mesh[c];
      T sx = (T)0, sy = (T)0, sz = (T)0;
#pragma unroll
      for (int a = 0; a < ORDER; ++a) {
        int64_t ix = i0[0] + a;
        if (ix >= n0) ix -= n0;
#pragma unroll
        for (int b = 0; b < ORDER; ++b) {
          int64_t iy = i0[1] + b;
          if (iy >= n1) iy -= n1;
          const T* row = m + (ix * n1 + iy) * n2;
#pragma unroll
          for (int e = 0; e < ORDER; ++e) {
            int64_t iz = i0[2] + e;
            if (iz >= n2) iz -= n2;
            const T v = row[iz];
            sx += v * dw[0][a] * w[1][b] * w[2][e];
            sy += v * w[0][a] * dw[1][b] * w[2][e];
            sz += v * w[0][a] * w[1][b] * dw[2][e];
          }
        }
      }
      g[0] += wt * sx;
      g[1] += wt * sy;
      g[2] += wt * sz;
    }
#pragma unroll
    for (int d = 0; d < 3; ++d) {
      const T v = G.scale * g[d] * P.inv_cell[d];
      G.gx[3 * p + d] = G.accumulate ? G.gx[3 * p + d] + v : v;
    }
  }
}

// acc (+)= M(k) in, with M the multiplier of fpm_kop (conjugated if conj).
template <typename T>
__global__ void k_kspace_accum(C<T>* acc, const C<T>* in, int64_t pitch2, int n0, int n1, int n2,
                               double L0, double L1, double L2, fpm_kop op, int conj,
                               int accumulate) {
  const int nh = n2 / 2 + 1;
  const int64_t total = (int64_t)n0 * n1 * nh;
  const double two_pi = 6.283185307179586476925286766559;
  const int nmv[3] = {n0, n1, n2};
  const double Lv[3] = {L0, L1, L2};
  for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < total;
       t += (int64_t)gridDim.x * blockDim.x) {
    const int k = (int)(t % nh);
    const int64_t r = t / nh;
    const int j = (int)(r % n1), i = (int)(r / n1);
    const int idx[3] = {i, j, k};
    double kk = 0.0, gfac = 1.0, wnd = 1.0;
#pragma unroll
    for (int d = 0; d < 3; ++d) {
      const int n = nmv[d];
      const bool half = d == 2;
      const int nn = half ? idx[d] : (idx[d] < (n + 1) / 2 ? idx[d] : idx[d] - n);
      const bool nyq = (n % 2 == 0) && (idx[d] == n / 2);
      const double kd = two_pi * nn / Lv[d];
      kk += kd * kd;
      const double h = Lv[d] / n;
      const double x = 2.0 * nn / n;
      const double g0 = nyq ? 0.0 : kd;
      const double g1 = nyq ? 0.0 : (8.0 * sinpi(x) - sinpi(2.0 * x)) / (6.0 * h);
      for (int g = 0; g < op.ngrad; ++g)
        if (op.grad_dir[g] == d) gfac *= op.grad_order == 0 ? g0 : g1;
      const double hx = (double)nn / n;
      const double sn = nn == 0 ? 1.0 : sinpi(hx) / (3.14159265358979323846 * hx);
      for (int q = 0; q < op.deconv; ++q) wnd /= sn;
    }
    double m = op.factor * gfac * wnd;
    if (op.laplace) m = kk == 0.0 ? 0.0 : m / kk;
    if (op.r_split != 0.0) m *= exp(-kk * op.r_split * op.r_split);
    // M = m i^ngrad; conj(M) = m (-i)^ngrad
    const int q = ((conj ? -op.ngrad : op.ngrad) % 4 + 4) % 4;
    const C<T> v = in[r * pitch2 + k];
    const double vx = v.x, vy = v.y;
    double ox, oy;
    switch (q) {
      case 0: ox = m * vx; oy = m * vy; break;
      case 1: ox = -m * vy; oy = m * vx; break;
      case 2: ox = -m * vx; oy = -m * vy; break;
      default: ox = m * vy; oy = -m * vx; break;
    }
    C<T>* a = acc + r * pitch2 + k;
    if (accumulate) {
      ox += (double)a->x;
      oy += (double)a->y;
    }
    *a = cmk<C<T>>((T)ox, (T)oy);
  }
}

template <typename T>
__global__ void k_hermitian_scale(C<T>* f, int64_t pitch2, int n0, int n1, int n2, T mid, T edge) {
  const int nh = n2 / 2 + 1;
  const int64_t total = (int64_t)n0 * n1 * nh;
  for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < total;
       t += (int64_t)gridDim.x * blockDim.x) {
    const int k = (int)(t % nh);
    const int64_t r = t / nh;
    const bool plane = k == 0 || ((n2 % 2 == 0) && k == n2 / 2);
    const T s = plane ? edge : mid;
    C<T>* p = f + r * pitch2 + k;
    *p = cmk<C<T>>(p->x * s, p->y * s);
  }
}

// source = 3/7 (xx yy + yy zz + zz xx - xy^2 - yz^2 - zx^2):
// d/dxx = 3/7 (yy + zz), ..., d/dxy = -6/7 xy, ...
template <typename T>
__global__ void k_lpt2_vjp(const T* gs, const T* xx, const T* yy, const T* zz, const T* xy,
                           const T* yz, const T* zx, T* gxx, T* gyy, T* gzz, T* gxy, T* gyz,
                           T* gzx, int64_t n) {
  const T c = (T)(3.0 / 7.0);
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x) {
    const T g = gs[i] * c;
    gxx[i] = g * (yy[i] + zz[i]);
    gyy[i] = g * (zz[i] + xx[i]);
    gzz[i] = g * (xx[i] + yy[i]);
    gxy[i] = (T)-2 * g * xy[i];
    gyz[i] = (T)-2 * g * yz[i];
    gzx[i] = (T)-2 * g * zx[i];
  }
}

template <typename T>
int fill_impl(T* p, int64_t n, T v, hipStream_t st);

template <typename T>
static int paint_w_impl(const fpm_particles* part, const fpm_mesh* m, int order, const void* wts,
                        int64_t wstride, double scale, void* out, int accumulate, hipStream_t st) {
  PartDesc<T> P;
  int rc = make_desc<T>(part, m, P);
  if (rc) return rc;
  if (!out || (P.n > 0 && !wts) || wstride < 1 || (order != 2 && order != 3)) return FPM_ERR_ARG;
  T* mesh = static_cast<T*>(out);
  const int n0 = m->nmesh[0], n1 = m->nmesh[1], n2 = m->nmesh[2];
  if (!accumulate && (rc = fill_impl<T>(mesh, (int64_t)n0 * n1 * n2, (T)0, st))) return rc;
  if (P.n == 0) return FPM_OK;
  const T* w = static_cast<const T*>(wts);
  if (order == 2)
    hipLaunchKernelGGL((k_paint_weighted<T, 2>), dim3(grid_n(P.n)), dim3(256), 0, st, P, mesh, n0,
                       n1, n2, w, wstride, (T)scale);
  else
    hipLaunchKernelGGL((k_paint_weighted<T, 3>), dim3(grid_n(P.n)), dim3(256), 0, st, P, mesh, n0,
                       n1, n2, w, wstride, (T)scale);
  FPM_LAUNCH_CHECK();
  return FPM_OK;
}

template <typename T>
static int window_grad_impl(const fpm_particles* part, const fpm_mesh* m, int order, int ncomp,
                            const void* const* meshes, const void* const* wts,
                            const int64_t* wstrides, double scale, void* gx, int accumulate,
                            hipStream_t st) {
  PartDesc<T> P;
  int rc = make_desc<T>(part, m, P);
  if (rc) return rc;
  if ((P.n > 0 && !gx) || !meshes || ncomp < 1 || ncomp > 4 || (order != 2 && order != 3))
    return FPM_ERR_ARG;
  GradArgs<T> G{};
  G.ncomp = ncomp;
  G.scale = (T)scale;
  G.gx = static_cast<T*>(gx);
  G.accumulate = accumulate ? 1 : 0;
  for (int c = 0; c < ncomp; ++c) {
    if (!meshes[c]) return FPM_ERR_ARG;
    G.mesh[c] = static_cast<const T*>(meshes[c]);
    G.wts[c] = wts ? static_cast<const T*>(wts[c]) : nullptr;
    G.wstride[c] = wstrides ? wstrides[c] : 1;
    if (G.wts[c] && G.wstride[c] < 1) return FPM_ERR_ARG;
  }
  if (P.n == 0) return FPM_OK;
  const int n0 = m->nmesh[0], n1 = m->nmesh[1], n2 = m->nmesh[2];
  if (order == 2)
    hipLaunchKernelGGL((k_window_grad<T, 2>), dim3(grid_n(P.n)), dim3(256), 0, st, P, G, n0, n1,
                       n2);
  else
    hipLaunchKernelGGL((k_window_grad<T, 3>), dim3(grid_n(P.n)), dim3(256), 0, st, P, G, n0, n1,
                       n2);
  FPM_LAUNCH_CHECK();
  return FPM_OK;
}

template <typename T>
static int accum_impl(void* acc, const void* in, const fpm_kop* op, int conj, int accumulate,
                      const fpm_mesh* m, int64_t pitch2, hipStream_t st) {
  if (!acc || !in || !op || !m || pitch2 < m->nmesh[2] / 2 + 1) return FPM_ERR_ARG;
  if (op->ngrad < 0 || op->ngrad > 2 || op->deconv < 0 || op->deconv > 8) return FPM_ERR_ARG;
  const int64_t total = (int64_t)m->nmesh[0] * m->nmesh[1] * (m->nmesh[2] / 2 + 1);
  hipLaunchKernelGGL(k_kspace_accum<T>, dim3(grid_n(total, 65536)), dim3(256), 0, st,
                     static_cast<C<T>*>(acc), static_cast<const C<T>*>(in), pitch2, m->nmesh[0],
                     m->nmesh[1], m->nmesh[2], m->box[0], m->box[1], m->box[2], *op, conj ? 1 : 0,
                     accumulate ? 1 : 0);
  FPM_LAUNCH_CHECK();
  return FPM_OK;
}

template <typename T>
static int herm_impl(void* f, double mid, double edge, const fpm_mesh* m, int64_t pitch2,
                     hipStream_t st) {
  if (!f || !m || pitch2 < m->nmesh[2] / 2 + 1) return FPM_ERR_ARG;
  const int64_t total = (int64_t)m->nmesh[0] * m->nmesh[1] * (m->nmesh[2] / 2 + 1);
  hipLaunchKernelGGL(k_hermitian_scale<T>, dim3(grid_n(total, 65536)), dim3(256), 0, st,
                     static_cast<C<T>*>(f), pitch2, m->nmesh[0], m->nmesh[1], m->nmesh[2], (T)mid,
                     (T)edge);
  FPM_LAUNCH_CHECK();
  return FPM_OK;
}

template <typename T>
static int lpt2_vjp_impl(const void* gs, const void* const* phi, void* const* gphi, int64_t n,
                         hipStream_t st) {
  if (!gs || !phi || !gphi || n < 0) return FPM_ERR_ARG;
  for (int i = 0; i < 6; ++i)
    if (!phi[i] || !gphi[i]) return FPM_ERR_ARG;
  if (n == 0) return FPM_OK;
  auto c = [](const void* p) { return static_cast<const T*>(p); };
  auto o = [](void* p) { return static_cast<T*>(p); };
  hipLaunchKernelGGL(k_lpt2_vjp<T>, dim3(grid_n(n, 65536)), dim3(256), 0, st, c(gs), c(phi[0]),
                     c(phi[1]), c(phi[2]), c(phi[3]), c(phi[4]), c(phi[5]), o(gphi[0]), o(gphi[1]),
                     o(gphi[2]), o(gphi[3]), o(gphi[4]), o(gphi[5]), n);
  FPM_LAUNCH_CHECK();
  return FPM_OK;
}

// ---------------------------------------------------------------- stencils
// The four-point gradient as a real-space stencil,
//   D_d f(x) = [8 (f(x+h) - f(x-h)) - (f(x+2h) - f(x-2h))] / (12 h),
// equal to the k-space multiplier i (8 sin kh - sin 2kh) / (6h) (the force's
// order-1 gradient) and antisymmetric (D_d^T = -D_d).  A workgroup per (row
// y, chunk of up to 256 z, kXChunk x planes) and a thread per z: each thread
// walks its x planes with the five values of its x stencil in registers, so
// the x-differenced mesh is loaded once per cell (four times with a
// workgroup per (x, y) row: 0.78 -> 0.54 ms for the divergence at 512^3);
// the y and z neighbours come through L1/L2; the only index arithmetic is
// one 32-bit division per workgroup and the periodic wraps.
__device__ __forceinline__ int wrap1(int j, int n) {
  j = j < 0 ? j + n : (j >= n ? j - n : j);
  return j < 0 ? j + n : (j >= n ? j - n : j);
}

template <typename T>
__device__ __forceinline__ T d4(const T* f, int64_t base, int i, int n, int64_t stride) {
  auto at = [&](int s) { return f[base + (int64_t)wrap1(i + s, n) * stride]; };
  return (T)8 * (at(1) - at(-1)) - (at(2) - at(-2));
}

constexpr int kXChunk = 16;

template <typename T>
__global__ __launch_bounds__(256) void k_grad4_div_x(const T* g0, const T* g1, const T* g2, T* out,
                                                     int n0, int n1, int n2, int nzc, T s0, T s1,
                                                     T s2, int accumulate) {
  const uint32_t b = blockIdx.x;
  const int y = (int)(b / (uint32_t)nzc);
  const int z = (int)(b - (uint32_t)y * (uint32_t)nzc) * (int)blockDim.x + (int)threadIdx.x;
  if (z >= n2) return;
  const int x0 = (int)blockIdx.y * kXChunk;
  const int x1 = x0 + kXChunk < n0 ? x0 + kXChunk : n0;
  const int64_t sx = (int64_t)n1 * n2, sy = n2;
  const int64_t yz = (int64_t)y * sy + z;
  T w[5];
#pragma unroll
  for (int k = 0; k < 4; ++k) w[k] = g0[(int64_t)wrap1(x0 - 2 + k, n0) * sx + yz];
  for (int x = x0; x < x1; ++x) {
    w[4] = g0[(int64_t)wrap1(x + 2, n0) * sx + yz];
    const int64_t c = (int64_t)x * sx + yz;
    const T dx = (T)8 * (w[3] - w[1]) - (w[4] - w[0]);
    const T v = s0 * dx + s1 * d4(g1, c - (int64_t)y * sy, y, n1, sy) +
                s2 * d4(g2, c - z, z, n2, (int64_t)1);
    out[c] = accumulate ? out[c] + v : v;
#pragma unroll
    for (int k = 0; k < 4; ++k) w[k] = w[k + 1];
  }
}

template <typename T>
__global__ __launch_bounds__(256) void k_grad4_mesh_x(const T* phi, T* o0, T* o1, T* o2, int n0,
                                                      int n1, int n2, int nzc, T s0, T s1, T s2) {
  const uint32_t b = blockIdx.x;
  const int y = (int)(b / (uint32_t)nzc);
  const int z = (int)(b - (uint32_t)y * (uint32_t)nzc) * (int)blockDim.x + (int)threadIdx.x;
  if (z >= n2) return;
  const int x0 = (int)blockIdx.y * kXChunk;
  const int x1 = x0 + kXChunk < n0 ? x0 + kXChunk : n0;
  const int64_t sx = (int64_t)n1 * n2, sy = n2;
  const int64_t yz = (int64_t)y * sy + z;
  T w[5];
#pragma unroll
  for (int k = 0; k < 4; ++k) w[k] = phi[(int64_t)wrap1(x0 - 2 + k, n0) * sx + yz];
  for (int x = x0; x < x1; ++x) {
    w[4] = phi[(int64_t)wrap1(x + 2, n0) * sx + yz];
    const int64_t c = (int64_t)x * sx + yz;
    o0[c] = s0 * ((T)8 * (w[3] - w[1]) - (w[4] - w[0]));
    o1[c] = s1 * d4(phi, c - (int64_t)y * sy, y, n1, sy);
    o2[c] = s2 * d4(phi, c - z, z, n2, (int64_t)1);
#pragma unroll
    for (int k = 0; k < 4; ++k) w[k] = w[k + 1];
  }
}

template <typename T>
static int grad4_mesh_impl(const void* phi, void* const* outs, const fpm_mesh* m, double scale,
                           hipStream_t st) {
  if (!phi || !outs || !m || !outs[0] || !outs[1] || !outs[2]) return FPM_ERR_ARG;
  for (int d = 0; d < 3; ++d)
    if (m->nmesh[d] <= 0 || !(m->box[d] > 0)) return FPM_ERR_ARG;
  T s[3];
  for (int d = 0; d < 3; ++d) s[d] = (T)(scale * m->nmesh[d] / (12.0 * m->box[d]));
  const unsigned nt = row_threads(m->nmesh[2]);
  const int nzc = (int)((m->nmesh[2] + nt - 1) / nt);
  const int64_t nb = (int64_t)m->nmesh[1] * nzc;
  const int nxc = (m->nmesh[0] + kXChunk - 1) / kXChunk;
  if (nb >= ((int64_t)1 << 31) || nxc > 65535) return FPM_ERR_UNSUPPORTED;
  hipLaunchKernelGGL(k_grad4_mesh_x<T>, dim3((unsigned)nb, (unsigned)nxc), dim3(nt), 0, st,
                     static_cast<const T*>(phi), static_cast<T*>(outs[0]),
                     static_cast<T*>(outs[1]), static_cast<T*>(outs[2]), m->nmesh[0],
                     m->nmesh[1], m->nmesh[2], nzc, s[0], s[1], s[2]);
  FPM_LAUNCH_CHECK();
  return FPM_OK;
}

template <typename T>
static int grad4_div_impl(const void* const* g, void* out, const fpm_mesh* m, double scale,
                          int accumulate, hipStream_t st) {
  if (!g || !out || !m || !g[0] || !g[1] || !g[2]) return FPM_ERR_ARG;
  for (int d = 0; d < 3; ++d)
    if (m->nmesh[d] <= 0 || !(m->box[d] > 0) || g[d] == out) return FPM_ERR_ARG;
  T s[3];
  for (int d = 0; d < 3; ++d) s[d] = (T)(scale * m->nmesh[d] / (12.0 * m->box[d]));
  const unsigned nt = row_threads(m->nmesh[2]);
  const int nzc = (int)((m->nmesh[2] + nt - 1) / nt);
  const int64_t nb = (int64_t)m->nmesh[1] * nzc;
  const int nxc = (m->nmesh[0] + kXChunk - 1) / kXChunk;
  if (nb >= ((int64_t)1 << 31) || nxc > 65535) return FPM_ERR_UNSUPPORTED;
  hipLaunchKernelGGL(k_grad4_div_x<T>, dim3((unsigned)nb, (unsigned)nxc), dim3(nt), 0, st,
                     static_cast<const T*>(g[0]), static_cast<const T*>(g[1]),
                     static_cast<const T*>(g[2]), static_cast<T*>(out), m->nmesh[0],
                     m->nmesh[1], m->nmesh[2], nzc, s[0], s[1], s[2], accumulate ? 1 : 0);
  FPM_LAUNCH_CHECK();
  return FPM_OK;
}

}  // namespace fpm

using namespace fpm;

extern "C" {

#define FPM_BOTH(NAME, IMPL, PARAMS, ARGS)                                      \
  int NAME##_f32 PARAMS { return IMPL<float> ARGS; }                           \
  int NAME##_f64 PARAMS { return IMPL<double> ARGS; }

FPM_BOTH(fpm_paint_weighted, paint_w_impl,
         (const fpm_particles* part, const fpm_mesh* mesh, int order, const void* weights,
          int64_t wstride, double scale, void* out, int accumulate, fpm_stream_t s),
         (part, mesh, order, weights, wstride, scale, out, accumulate, (hipStream_t)s))

FPM_BOTH(fpm_window_grad, window_grad_impl,
         (const fpm_particles* part, const fpm_mesh* mesh, int order, int ncomp,
          const void* const* meshes, const void* const* weights, const int64_t* wstrides,
          double scale, void* gx, int accumulate, fpm_stream_t s),
         (part, mesh, order, ncomp, meshes, weights, wstrides, scale, gx, accumulate,
          (hipStream_t)s))

FPM_BOTH(fpm_kspace_accumulate, accum_impl,
         (void* acc, const void* in, const fpm_kop* op, int conj, int accumulate,
          const fpm_mesh* mesh, int64_t pitch2, fpm_stream_t s),
         (acc, in, op, conj, accumulate, mesh, pitch2, (hipStream_t)s))

FPM_BOTH(fpm_hermitian_scale, herm_impl,
         (void* field, double mid, double edge, const fpm_mesh* mesh, int64_t pitch2,
          fpm_stream_t s),
         (field, mid, edge, mesh, pitch2, (hipStream_t)s))

FPM_BOTH(fpm_lpt2_source_vjp, lpt2_vjp_impl,
         (const void* gsource, const void* const* phi, void* const* gphi, int64_t ncells,
          fpm_stream_t s),
         (gsource, phi, gphi, ncells, (hipStream_t)s))

FPM_BOTH(fpm_grad4_mesh, grad4_mesh_impl,
         (const void* phi, void* const* outs, const fpm_mesh* mesh, double scale, fpm_stream_t s),
         (phi, outs, mesh, scale, (hipStream_t)s))

FPM_BOTH(fpm_grad4_div, grad4_div_impl,
         (const void* const* g, void* out, const fpm_mesh* mesh, double scale, int accumulate,
          fpm_stream_t s),
         (g, out, mesh, scale, accumulate, (hipStream_t)s))

#undef FPM_BOTH

}  // extern "C"
