// Particle-array updates: kick / drift (axpy with a double coefficient), the
// 2LPT combination of displacement, momentum and force, and positions X = Q + S.
#include "fpm_common.h"

namespace fpm {

template <typename T>
int make_desc(const fpm_particles* part, const fpm_mesh* m, PartDesc<T>& P);

static unsigned grid_elems(int64_t n) {
  int64_t g = (n + 255) / 256;
  if (g < 1) g = 1;
  if (g > 65536) g = 65536;
  return (unsigned)g;
}

template <typename T>
__global__ void k_axpy(T* y, const T* x, int64_t n, T a) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    y[i] = madd(a, x[i], y[i]);
}

// nkick kicks then a drift in one pass: P += k0 F, P += k1 F, S += d P, each
// rounded as k_axpy rounds it, with P, F and S read once.
template <typename T>
__global__ void k_kick_drift(T* S, T* P, const T* F, int64_t n, T k0, T k1, int nkick, T d) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x) {
    T p = P[i];
    if (nkick > 0) {
      const T f = F[i];
      p = madd(k0, f, p);
      if (nkick > 1) p = madd(k1, f, p);
      P[i] = p;
    }
    S[i] = madd(d, p, S[i]);
  }
}

// The same on four-element vectors (float32 arrays with 16-byte aligned
// pointers): a quarter of the load and store instructions, the same
// arithmetic per element.  The n % 4 tail elements go to the first threads.
__global__ void k_kick_drift_v4(float* S, float* P, const float* F, int64_t n, float k0,
                                float k1, int nkick, float d) {
  typedef float V4 __attribute__((ext_vector_type(4)));
  const int64_t n4 = n / 4;
  V4* S4 = reinterpret_cast<V4*>(S);
  V4* P4 = reinterpret_cast<V4*>(P);
  const V4* F4 = reinterpret_cast<const V4*>(F);
  const int64_t tid = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  for (int64_t i = tid; i < n4; i += (int64_t)gridDim.x * blockDim.x) {
    V4 p = P4[i];
    if (nkick > 0) {
      const V4 f = F4[i];
#pragma unroll
      for (int c = 0; c < 4; ++c) {
        p[c] = madd(k0, f[c], p[c]);
        if (nkick > 1) p[c] = madd(k1, f[c], p[c]);
      }
      P4[i] = p;
    }
    V4 x = S4[i];
#pragma unroll
    for (int c = 0; c < 4; ++c) x[c] = madd(d, p[c], x[c]);
    S4[i] = x;
  }
  const int64_t r = n4 * 4 + tid;
  if (r < n) {
    float p = P[r];
    if (nkick > 0) {
      const float f = F[r];
      p = madd(k0, f, p);
      if (nkick > 1) p = madd(k1, f, p);
      P[r] = p;
    }
    S[r] = madd(d, p, S[r]);
  }
}

// Adjoint of a step boundary, traversed backwards: the drift's transpose
// gP += d gS, then the kicks' transposes gF += k_i gP, each rounded as
// k_axpy.  fresh = 1: gF holds no gradient yet (it was consumed by a force's
// transpose), so it starts from zero and is not read.  With no kick gF is
// left alone.
template <typename T>
__global__ void k_drift_kick_vjp(const T* gS, T* gP, T* gF, int64_t n, T d, T k0, T k1,
                                 int nkick, int fresh) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x) {
    const T p = madd(d, gS[i], gP[i]);
    gP[i] = p;
    if (nkick > 0) {
      T f = fresh ? (T)0 : gF[i];
      f = madd(k0, p, f);
      if (nkick > 1) f = madd(k1, p, f);
      gF[i] = f;
    }
  }
}

// y += table[counter[0] * stride + slot] * x: the coefficient is read on the
// device, so one captured graph can replay every step of a run.  A counter
// outside [0, rows) leaves y unchanged.
template <typename T>
__global__ void k_axpy_table(T* y, const T* x, int64_t n, const double* table,
                             const int32_t* counter, int rows, int stride, int slot) {
  const int32_t row = counter[0];
  if (row < 0 || row >= rows) return;
  const T a = (T)table[(int64_t)row * stride + slot];
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    y[i] = madd(a, x[i], y[i]);
}

// Table-driven step boundary for a captured graph: with row = counter[0],
// P += table[row-1][close] F (the previous step's closing kick, row > 0 only),
// P += table[row][open] F, S += table[row][drift] P, each rounded as k_axpy.
// A counter outside [0, rows) leaves S and P unchanged.
template <typename T>
__global__ void k_kick_drift_table(T* S, T* P, const T* F, int64_t n, const double* table,
                                   const int32_t* counter, int rows, int stride, int close,
                                   int open, int drift) {
  const int32_t row = counter[0];
  if (row < 0 || row >= rows) return;
  const bool prev = row > 0;
  const T k0 = prev ? (T)table[(int64_t)(row - 1) * stride + close] : (T)0;
  const T k1 = (T)table[(int64_t)row * stride + open];
  const T d = (T)table[(int64_t)row * stride + drift];
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x) {
    const T f = F[i];
    T p = P[i];
    if (prev) p = madd(k0, f, p);
    p = madd(k1, f, p);
    P[i] = p;
    S[i] = madd(d, p, S[i]);
  }
}

// y = a y + b x (the damped kick of the glass generator).
template <typename T>
__global__ void k_axpby(T* y, const T* x, int64_t n, T a, T b) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    y[i] = a * y[i] + b * x[i];
}

// s = (s + a p) mod L_d on [N, 3] arrays, the result in [0, L_d).
template <typename T>
__global__ void k_drift_wrap(T* s, const T* p, int64_t n3, T a, T L0, T L1, T L2) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n3;
       i += (int64_t)gridDim.x * blockDim.x) {
    const int d = (int)(i % 3);
    const T L = d == 0 ? L0 : (d == 1 ? L1 : L2);
    T v = s[i] + a * p[i];
    v = v - L * floor(v / L);
    s[i] = v < L ? v : v - L;  // v / L rounding up to 1 leaves v == L
  }
}

__global__ void k_counter_add(int32_t* counter, int32_t v) {
  if (threadIdx.x == 0 && blockIdx.x == 0) counter[0] += v;
}

template <typename T>
__global__ void k_lpt_combine(const T* dx1, const T* dx2, T* S, T* P, T* F, int64_t n, T c0, T c1,
                              T c2, T c3, T c4, T c5) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x) {
    const T a = dx1[i];
    const T b = dx2 ? dx2[i] : (T)0;
    S[i] = c0 * a + c1 * b;
    P[i] = c2 * a + c3 * b;
    F[i] = c4 * a + c5 * b;
  }
}

// Positions in box units.  For the lattice the integer lattice coordinate is
// converted once; X = (i + shift) * L / Nlat + S.
template <typename T>
__global__ void k_positions(PartDesc<T> P, double sx, double sy, double sz, T* out) {
  const double sp[3] = {sx, sy, sz};
  for (int64_t p = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; p < P.n;
       p += (int64_t)gridDim.x * blockDim.x) {
    if (P.q == nullptr) {
      const int64_t iz = p % P.lat[2];
      const int64_t t = p / P.lat[2];
      const int64_t li[3] = {t / P.lat[1], t % P.lat[1], iz};
#pragma unroll
      for (int d = 0; d < 3; ++d) {
        const T q = (T)(((double)li[d] + P.lat_shift) * sp[d]);
        out[3 * p + d] = q + (P.s ? P.s[3 * p + d] : (T)0);
      }
    } else {
#pragma unroll
      for (int d = 0; d < 3; ++d) out[3 * p + d] = P.q[3 * p + d] + (P.s ? P.s[3 * p + d] : (T)0);
    }
  }
}

template <typename T>
static int axpy_impl(void* y, const void* x, int64_t n, double a, hipStream_t st) {
  if (n < 0 || (n > 0 && (!y || !x))) return FPM_ERR_ARG;
  if (n == 0) return FPM_OK;
  hipLaunchKernelGGL(k_axpy<T>, dim3(grid_elems(n)), dim3(256), 0, st, static_cast<T*>(y),
                     static_cast<const T*>(x), n, (T)a);
  FPM_LAUNCH_CHECK();
  return FPM_OK;
}

template <typename T>
static int kick_drift_impl(void* S, void* P, const void* F, int64_t n, const double* kicks,
                           int nkick, double drift, hipStream_t st) {
  if (n < 0 || nkick < 0 || nkick > 2 || (nkick > 0 && !kicks) ||
      (n > 0 && (!S || !P || (nkick > 0 && !F))))
    return FPM_ERR_ARG;
  if (n == 0) return FPM_OK;
  const T k0 = nkick > 0 ? (T)kicks[0] : (T)0, k1 = nkick > 1 ? (T)kicks[1] : (T)0;
  auto aligned = [](const void* q) { return q == nullptr || ((uintptr_t)q & 15) == 0; };
  if (sizeof(T) == 4 && aligned(S) && aligned(P) && aligned(F)) {
    hipLaunchKernelGGL(k_kick_drift_v4, dim3(grid_elems((n + 3) / 4)), dim3(256), 0, st,
                       static_cast<float*>(S), static_cast<float*>(P),
                       static_cast<const float*>(F), n, (float)k0, (float)k1, nkick,
                       (float)drift);
  } else {
    hipLaunchKernelGGL(k_kick_drift<T>, dim3(grid_elems(n)), dim3(256), 0, st,
                       static_cast<T*>(S), static_cast<T*>(P), static_cast<const T*>(F), n, k0,
                       k1, nkick, (T)drift);
  }
  FPM_LAUNCH_CHECK();
  return FPM_OK;
}

template <typename T>
static int drift_kick_vjp_impl(const void* gS, void* gP, void* gF, int64_t n, const double* kicks,
                               int nkick, double drift, int fresh, hipStream_t st) {
  if (n < 0 || nkick < 0 || nkick > 2 || (nkick > 0 && !kicks) ||
      (n > 0 && (!gS || !gP || (nkick > 0 && !gF))))
    return FPM_ERR_ARG;
  if (n == 0) return FPM_OK;
  const T k0 = nkick > 0 ? (T)kicks[0] : (T)0, k1 = nkick > 1 ? (T)kicks[1] : (T)0;
  hipLaunchKernelGGL(k_drift_kick_vjp<T>, dim3(grid_elems(n)), dim3(256), 0, st,
                     static_cast<const T*>(gS), static_cast<T*>(gP), static_cast<T*>(gF), n,
                     (T)drift, k0, k1, nkick, fresh ? 1 : 0);
  FPM_LAUNCH_CHECK();
  return FPM_OK;
}

template <typename T>
static int axpy_table_impl(void* y, const void* x, int64_t n, const double* table,
                           const int32_t* counter, int rows, int stride, int slot,
                           hipStream_t st) {
  if (n < 0 || !table || !counter || rows < 1 || stride < 1 || slot < 0 || slot >= stride ||
      (n > 0 && (!y || !x)))
    return FPM_ERR_ARG;
  if (n == 0) return FPM_OK;
  hipLaunchKernelGGL(k_axpy_table<T>, dim3(grid_elems(n)), dim3(256), 0, st, static_cast<T*>(y),
                     static_cast<const T*>(x), n, table, counter, rows, stride, slot);
  FPM_LAUNCH_CHECK();
  return FPM_OK;
}

template <typename T>
static int kick_drift_table_impl(void* S, void* P, const void* F, int64_t n, const double* table,
                                 const int32_t* counter, int rows, int stride, int close,
                                 int open, int drift, hipStream_t st) {
  auto bad = [stride](int s) { return s < 0 || s >= stride; };
  if (n < 0 || !table || !counter || rows < 1 || stride < 1 || bad(close) || bad(open) ||
      bad(drift) || (n > 0 && (!S || !P || !F)))
    return FPM_ERR_ARG;
  if (n == 0) return FPM_OK;
  hipLaunchKernelGGL(k_kick_drift_table<T>, dim3(grid_elems(n)), dim3(256), 0, st,
                     static_cast<T*>(S), static_cast<T*>(P), static_cast<const T*>(F), n, table,
                     counter, rows, stride, close, open, drift);
  FPM_LAUNCH_CHECK();
  return FPM_OK;
}

template <typename T>
static int axpby_impl(void* y, const void* x, int64_t n, double a, double b, hipStream_t st) {
  if (n < 0 || (n > 0 && (!y || !x))) return FPM_ERR_ARG;
  if (n == 0) return FPM_OK;
  hipLaunchKernelGGL(k_axpby<T>, dim3(grid_elems(n)), dim3(256), 0, st, static_cast<T*>(y),
                     static_cast<const T*>(x), n, (T)a, (T)b);
  FPM_LAUNCH_CHECK();
  return FPM_OK;
}

template <typename T>
static int drift_wrap_impl(void* s, const void* p, int64_t n, double a, const double* box,
                           hipStream_t st) {
  if (n < 0 || !box || (n > 0 && (!s || !p))) return FPM_ERR_ARG;
  for (int d = 0; d < 3; ++d)
    if (!(box[d] > 0)) return FPM_ERR_ARG;
  if (n == 0) return FPM_OK;
  hipLaunchKernelGGL(k_drift_wrap<T>, dim3(grid_elems(3 * n)), dim3(256), 0, st,
                     static_cast<T*>(s), static_cast<const T*>(p), 3 * n, (T)a, (T)box[0],
                     (T)box[1], (T)box[2]);
  FPM_LAUNCH_CHECK();
  return FPM_OK;
}

template <typename T>
static int combine_impl(const void* dx1, const void* dx2, void* S, void* P, void* F, int64_t n,
                        const double* c, hipStream_t st) {
  if (n < 0 || !c || (n > 0 && (!dx1 || !S || !P || !F))) return FPM_ERR_ARG;
  if (n == 0) return FPM_OK;
  hipLaunchKernelGGL(k_lpt_combine<T>, dim3(grid_elems(n)), dim3(256), 0, st,
                     static_cast<const T*>(dx1), static_cast<const T*>(dx2), static_cast<T*>(S),
                     static_cast<T*>(P), static_cast<T*>(F), n, (T)c[0], (T)c[1], (T)c[2],
                     (T)c[3], (T)c[4], (T)c[5]);
  FPM_LAUNCH_CHECK();
  return FPM_OK;
}

template <typename T>
static int positions_impl(const fpm_particles* part, const fpm_mesh* m, void* out,
                          hipStream_t st) {
  PartDesc<T> P;
  int rc = make_desc<T>(part, m, P);
  if (rc) return rc;
  if (P.n == 0) return FPM_OK;
  if (!out) return FPM_ERR_ARG;
  double sp[3] = {1, 1, 1};
  if (!part->q)
    for (int d = 0; d < 3; ++d) sp[d] = m->box[d] / (double)part->lattice[d];
  hipLaunchKernelGGL(k_positions<T>, dim3(grid_elems(P.n)), dim3(256), 0, st, P, sp[0], sp[1],
                     sp[2], static_cast<T*>(out));
  FPM_LAUNCH_CHECK();
  return FPM_OK;
}

}  // namespace fpm

using namespace fpm;

extern "C" {

int fpm_axpy_f32(void* y, const void* x, int64_t n, double a, fpm_stream_t s) {
  return axpy_impl<float>(y, x, n, a, (hipStream_t)s);
}
int fpm_axpy_f64(void* y, const void* x, int64_t n, double a, fpm_stream_t s) {
  return axpy_impl<double>(y, x, n, a, (hipStream_t)s);
}
int fpm_kick_drift_f32(void* S, void* P, const void* F, int64_t n, const double* kicks,
                      int nkick, double drift, fpm_stream_t s) {
  return kick_drift_impl<float>(S, P, F, n, kicks, nkick, drift, (hipStream_t)s);
}
int fpm_kick_drift_f64(void* S, void* P, const void* F, int64_t n, const double* kicks,
                      int nkick, double drift, fpm_stream_t s) {
  return kick_drift_impl<double>(S, P, F, n, kicks, nkick, drift, (hipStream_t)s);
}
int fpm_drift_kick_vjp_f32(const void* gS, void* gP, void* gF, int64_t n, const double* kicks,
                           int nkick, double drift, int fresh, fpm_stream_t s) {
  return drift_kick_vjp_impl<float>(gS, gP, gF, n, kicks, nkick, drift, fresh, (hipStream_t)s);
}
int fpm_drift_kick_vjp_f64(const void* gS, void* gP, void* gF, int64_t n, const double* kicks,
                           int nkick, double drift, int fresh, fpm_stream_t s) {
  return drift_kick_vjp_impl<double>(gS, gP, gF, n, kicks, nkick, drift, fresh, (hipStream_t)s);
}
int fpm_axpy_table_f32(void* y, const void* x, int64_t n, const double* table,
                       const int32_t* counter, int rows, int stride, int slot, fpm_stream_t s) {
  return axpy_table_impl<float>(y, x, n, table, counter, rows, stride, slot, (hipStream_t)s);
}
int fpm_axpy_table_f64(void* y, const void* x, int64_t n, const double* table,
                       const int32_t* counter, int rows, int stride, int slot, fpm_stream_t s) {
  return axpy_table_impl<double>(y, x, n, table, counter, rows, stride, slot, (hipStream_t)s);
}
int fpm_kick_drift_table_f32(void* S, void* P, const void* F, int64_t n, const double* table,
                             const int32_t* counter, int rows, int stride, int close, int open,
                             int drift, fpm_stream_t s) {
  return kick_drift_table_impl<float>(S, P, F, n, table, counter, rows, stride, close, open, drift,
                                      (hipStream_t)s);
}
int fpm_kick_drift_table_f64(void* S, void* P, const void* F, int64_t n, const double* table,
                             const int32_t* counter, int rows, int stride, int close, int open,
                             int drift, fpm_stream_t s) {
  return kick_drift_table_impl<double>(S, P, F, n, table, counter, rows, stride, close, open,
                                       drift, (hipStream_t)s);
}
int fpm_axpby_f32(void* y, const void* x, int64_t n, double a, double b, fpm_stream_t s) {
  return axpby_impl<float>(y, x, n, a, b, (hipStream_t)s);
}
int fpm_axpby_f64(void* y, const void* x, int64_t n, double a, double b, fpm_stream_t s) {
  return axpby_impl<double>(y, x, n, a, b, (hipStream_t)s);
}
int fpm_drift_wrap_f32(void* S, const void* P, int64_t n, double a, const double* box,
                       fpm_stream_t s) {
  return drift_wrap_impl<float>(S, P, n, a, box, (hipStream_t)s);
}
int fpm_drift_wrap_f64(void* S, const void* P, int64_t n, double a, const double* box,
                       fpm_stream_t s) {
  return drift_wrap_impl<double>(S, P, n, a, box, (hipStream_t)s);
}
int fpm_counter_add(int32_t* counter, int32_t v, fpm_stream_t s) {
  if (!counter) return FPM_ERR_ARG;
  hipLaunchKernelGGL(k_counter_add, dim3(1), dim3(64), 0, (hipStream_t)s, counter, v);
  FPM_LAUNCH_CHECK();
  return FPM_OK;
}
int fpm_lpt_combine_f32(const void* dx1, const void* dx2, void* S, void* P, void* F, int64_t n,
                        const double* c, fpm_stream_t s) {
  return combine_impl<float>(dx1, dx2, S, P, F, n, c, (hipStream_t)s);
}
int fpm_lpt_combine_f64(const void* dx1, const void* dx2, void* S, void* P, void* F, int64_t n,
                        const double* c, fpm_stream_t s) {
  return combine_impl<double>(dx1, dx2, S, P, F, n, c, (hipStream_t)s);
}
int fpm_positions_f32(const fpm_particles* part, const fpm_mesh* mesh, void* out, fpm_stream_t s) {
  return positions_impl<float>(part, mesh, out, (hipStream_t)s);
}
int fpm_positions_f64(const fpm_particles* part, const fpm_mesh* mesh, void* out, fpm_stream_t s) {
  return positions_impl<double>(part, mesh, out, (hipStream_t)s);
}

}  // extern "C"
