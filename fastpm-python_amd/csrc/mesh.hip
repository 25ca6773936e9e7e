// Mass assignment (paint) and window interpolation (readout) on periodic meshes.
//
// Windows, in mesh-cell units with the particle at x = cell + f, f in [0, 1):
//   CIC (order 2): points cell, cell+1 with weights 1-f, f.
//   TSC (order 3): nearest point c = cell + (f >= 1/2), d = x - c in [-1/2, 1/2);
//                  points c-1, c, c+1 with 1/2 (1/2-d)^2, 3/4 - d^2, 1/2 (1/2+d)^2.
// A particle on a grid point deposits everything there (CIC) or 1/8, 3/4, 1/8
// per axis (TSC).  Paint scatters with hardware float atomics (global_atomic_add,
// no return value); readout gathers all requested meshes in one pass.
#include "fpm_common.h"

namespace fpm {

template <typename T, int ORDER>
struct Window {
  int64_t i0[3];
  T w[3][ORDER];
};

template <typename T, int ORDER>
__device__ __forceinline__ void window(const int64_t cell[3], const T f[3], const int nm[3],
                                       Window<T, ORDER>& W) {
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    if constexpr (ORDER == 2) {
      W.i0[d] = cell[d];
      W.w[d][0] = (T)1 - f[d];
      W.w[d][1] = f[d];
    } else {
      const bool up = f[d] >= (T)0.5;
      const T dd = up ? f[d] - (T)1 : f[d];
      W.i0[d] = cell[d] + (up ? 1 : 0) - 1;
      const T a = (T)0.5 - dd, b = (T)0.5 + dd;
      W.w[d][0] = (T)0.5 * a * a;
      W.w[d][1] = (T)0.75 - dd * dd;
      W.w[d][2] = (T)0.5 * b * b;
    }
    W.i0[d] = wrapi(W.i0[d], nm[d]);
  }
}

template <typename T>
__global__ void k_fill(T* p, int64_t n, T v) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    p[i] = v;
}

template <typename T, int ORDER>
__global__ __launch_bounds__(256) void k_paint(PartDesc<T> P, T* mesh, int n0, int n1, int n2,
                                                T mass) {
  const int nm[3] = {n0, n1, n2};
  for (int64_t p = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; p < P.n;
       p += (int64_t)gridDim.x * blockDim.x) {
    int64_t cell[3];
    T f[3];
    particle_cell(P, p, cell, f);
    Window<T, ORDER> W;
    window<T, ORDER>(cell, f, nm, W);
#pragma unroll
    for (int a = 0; a < ORDER; ++a) {
      int64_t ix = W.i0[0] + a;
      if (ix >= n0) ix -= n0;
      const T wa = W.w[0][a] * mass;
#pragma unroll
      for (int b = 0; b < ORDER; ++b) {
        int64_t iy = W.i0[1] + b;
        if (iy >= n1) iy -= n1;
        const T wab = wa * W.w[1][b];
        T* row = mesh + (ix * n1 + iy) * n2;
#pragma unroll
        for (int c = 0; c < ORDER; ++c) {
          int64_t iz = W.i0[2] + c;
          if (iz >= n2) iz -= n2;
          atomicAdd(row + iz, wab * W.w[2][c]);
        }
      }
    }
  }
}

template <typename T>
struct ReadArgs {
  const T* mesh[4];
  T* out[4];
  int64_t stride[4];
  T scale[4];
  int ncomp;
};

template <typename T, int ORDER, int NC>
__global__ __launch_bounds__(256) void k_readout(PartDesc<T> P, ReadArgs<T> R, int n0, int n1,
                                                  int n2) {
  const int nm[3] = {n0, n1, n2};
  for (int64_t p = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; p < P.n;
       p += (int64_t)gridDim.x * blockDim.x) {
    int64_t cell[3];
    T f[3];
    particle_cell(P, p, cell, f);
    Window<T, ORDER> W;
    window<T, ORDER>(cell, f, nm, W);
    T acc[NC];
#pragma unroll
    for (int c = 0; c < NC; ++c) acc[c] = (T)0;
#pragma unroll
    for (int a = 0; a < ORDER; ++a) {
      int64_t ix = W.i0[0] + a;
      if (ix >= n0) ix -= n0;
#pragma unroll
      for (int b = 0; b < ORDER; ++b) {
        int64_t iy = W.i0[1] + b;
        if (iy >= n1) iy -= n1;
        const T wab = W.w[0][a] * W.w[1][b];
        const int64_t row = (ix * n1 + iy) * n2;
#pragma unroll
        for (int c = 0; c < ORDER; ++c) {
          int64_t iz = W.i0[2] + c;
          if (iz >= n2) iz -= n2;
          const T w = wab * W.w[2][c];
#pragma unroll
          for (int q = 0; q < NC; ++q) acc[q] += w * R.mesh[q][row + iz];
        }
      }
    }
#pragma unroll
    for (int q = 0; q < NC; ++q) R.out[q][p * R.stride[q]] = acc[q] * R.scale[q];
  }
}

template <typename T>
int make_desc(const fpm_particles* part, const fpm_mesh* m, PartDesc<T>& P) {
  if (!part || !m || part->n < 0) return FPM_ERR_ARG;
  P.lat_shift = 0.0;
  if (part->n == 0) {
    // no particles: no arrays needed (every entry point returns after its
    // mesh outputs are written)
    P.q = nullptr;
    P.s = nullptr;
    P.n = 0;
    for (int d = 0; d < 3; ++d) {
      if (m->nmesh[d] <= 0 || !(m->box[d] > 0)) return FPM_ERR_ARG;
      P.lat[d] = 0;
      P.ratio[d] = 1.0;
      P.inv_cell[d] = (T)((double)m->nmesh[d] / m->box[d]);
    }
    return FPM_OK;
  }
  P.q = static_cast<const T*>(part->q);
  P.s = static_cast<const T*>(part->s);
  P.n = part->n;
  P.lat_shift = part->lattice_shift;
  int64_t lat_total = 1;
  for (int d = 0; d < 3; ++d) {
    if (m->nmesh[d] <= 0 || !(m->box[d] > 0)) return FPM_ERR_ARG;
    P.lat[d] = part->lattice[d];
    if (!part->q) {
      if (part->lattice[d] <= 0) return FPM_ERR_ARG;
      lat_total *= part->lattice[d];
      P.ratio[d] = (double)m->nmesh[d] / (double)part->lattice[d];
    } else {
      P.ratio[d] = 1.0;
    }
    P.inv_cell[d] = (T)((double)m->nmesh[d] / m->box[d]);
  }
  if (!part->q && lat_total != part->n) return FPM_ERR_ARG;
  return FPM_OK;
}
template int make_desc<float>(const fpm_particles*, const fpm_mesh*, PartDesc<float>&);
template int make_desc<double>(const fpm_particles*, const fpm_mesh*, PartDesc<double>&);

static unsigned grid_for(int64_t n, int64_t cap = 1 << 20) {
  int64_t g = (n + 255) / 256;
  if (g < 1) g = 1;
  if (g > cap) g = cap;
  return (unsigned)g;
}

template <typename T>
int fill_impl(T* p, int64_t n, T v, hipStream_t st) {
  if (n <= 0) return FPM_OK;
  hipLaunchKernelGGL(k_fill<T>, dim3(grid_for(n, 65536)), dim3(256), 0, st, p, n, v);
  FPM_LAUNCH_CHECK();
  return FPM_OK;
}
template int fill_impl<float>(float*, int64_t, float, hipStream_t);
template int fill_impl<double>(double*, int64_t, double, hipStream_t);

template <typename T>
static int paint_impl(const fpm_particles* part, const fpm_mesh* m, int order, double mass,
                      void* out, int accumulate, hipStream_t st) {
  PartDesc<T> P;
  int rc = make_desc<T>(part, m, P);
  if (rc) return rc;
  if (!out || (order != 2 && order != 3)) return FPM_ERR_ARG;
  T* mesh = static_cast<T*>(out);
  const int n0 = m->nmesh[0], n1 = m->nmesh[1], n2 = m->nmesh[2];
  if (!accumulate && (rc = fill_impl<T>(mesh, (int64_t)n0 * n1 * n2, (T)0, st))) return rc;
  if (P.n == 0) return FPM_OK;
  if (order == 2)
    hipLaunchKernelGGL((k_paint<T, 2>), dim3(grid_for(P.n)), dim3(256), 0, st, P, mesh, n0, n1, n2,
                       (T)mass);
  else
    hipLaunchKernelGGL((k_paint<T, 3>), dim3(grid_for(P.n)), dim3(256), 0, st, P, mesh, n0, n1, n2,
                       (T)mass);
  FPM_LAUNCH_CHECK();
  return FPM_OK;
}

template <typename T, int ORDER>
static void launch_readout(const PartDesc<T>& P, const ReadArgs<T>& R, const fpm_mesh* m,
                           hipStream_t st) {
  const int n0 = m->nmesh[0], n1 = m->nmesh[1], n2 = m->nmesh[2];
  const dim3 g(grid_for(P.n)), b(256);
  switch (R.ncomp) {
    case 1: hipLaunchKernelGGL((k_readout<T, ORDER, 1>), g, b, 0, st, P, R, n0, n1, n2); break;
    case 2: hipLaunchKernelGGL((k_readout<T, ORDER, 2>), g, b, 0, st, P, R, n0, n1, n2); break;
    case 3: hipLaunchKernelGGL((k_readout<T, ORDER, 3>), g, b, 0, st, P, R, n0, n1, n2); break;
    default: hipLaunchKernelGGL((k_readout<T, ORDER, 4>), g, b, 0, st, P, R, n0, n1, n2); break;
  }
}

template <typename T>
static int readout_impl(const fpm_particles* part, const fpm_mesh* m, int order, int ncomp,
                        const void* const* meshes, void* const* outs, const int64_t* strides,
                        const double* scales, hipStream_t st) {
  PartDesc<T> P;
  int rc = make_desc<T>(part, m, P);
  if (rc) return rc;
  if ((order != 2 && order != 3) || ncomp < 1 || ncomp > 4 || !meshes || !outs || !strides)
    return FPM_ERR_ARG;
  ReadArgs<T> R{};
  R.ncomp = ncomp;
  for (int c = 0; c < ncomp; ++c) {
    if (!meshes[c] || (P.n > 0 && !outs[c]) || strides[c] < 1) return FPM_ERR_ARG;
    R.mesh[c] = static_cast<const T*>(meshes[c]);
    R.out[c] = static_cast<T*>(outs[c]);
    R.stride[c] = strides[c];
    R.scale[c] = scales ? (T)scales[c] : (T)1;
  }
  if (P.n == 0) return FPM_OK;
  if (order == 2) launch_readout<T, 2>(P, R, m, st);
  else launch_readout<T, 3>(P, R, m, st);
  FPM_LAUNCH_CHECK();
  return FPM_OK;
}

}  // namespace fpm

using namespace fpm;

extern "C" {

int fpm_paint_f32(const fpm_particles* part, const fpm_mesh* mesh, int order, double mass,
                  void* out, int accumulate, fpm_stream_t s) {
  return paint_impl<float>(part, mesh, order, mass, out, accumulate, (hipStream_t)s);
}
int fpm_paint_f64(const fpm_particles* part, const fpm_mesh* mesh, int order, double mass,
                  void* out, int accumulate, fpm_stream_t s) {
  return paint_impl<double>(part, mesh, order, mass, out, accumulate, (hipStream_t)s);
}
int fpm_readout_f32(const fpm_particles* part, const fpm_mesh* mesh, int order, int ncomp,
                    const void* const* meshes, void* const* outs, const int64_t* strides,
                    const double* scales, fpm_stream_t s) {
  return readout_impl<float>(part, mesh, order, ncomp, meshes, outs, strides, scales,
                             (hipStream_t)s);
}
int fpm_readout_f64(const fpm_particles* part, const fpm_mesh* mesh, int order, int ncomp,
                    const void* const* meshes, void* const* outs, const int64_t* strides,
                    const double* scales, fpm_stream_t s) {
  return readout_impl<double>(part, mesh, order, ncomp, meshes, outs, strides, scales,
                              (hipStream_t)s);
}

}  // extern "C"
