// Spectral-side kernels: white noise, linear-power scaling, the 2LPT source
// product, and double-precision shell-binned spectra.
#include <mutex>

#include "fpm_common.h"

namespace fpm {

static unsigned grid_elems(int64_t n) {
  int64_t g = (n + 255) / 256;
  if (g < 1) g = 1;
  if (g > 65536) g = 65536;
  return (unsigned)g;
}

// ---------------------------------------------------------------- Philox4x32-10
// Counter-based generator of Salmon, Moraes, Dror and Shaw, "Parallel random
// numbers: as easy as 1, 2, 3" (SC11): ten rounds of two 32x32->64 multiplies,
// with the Weyl key schedule between rounds.
struct U4 {
  uint32_t v[4];
};

__device__ __forceinline__ U4 philox4x32_10(U4 c, uint32_t k0, uint32_t k1) {
  const uint32_t M0 = 0xD2511F53u, M1 = 0xCD9E8D57u;
  const uint32_t W0 = 0x9E3779B9u, W1 = 0xBB67AE85u;
#pragma unroll
  for (int r = 0; r < 10; ++r) {
    const uint64_t p0 = (uint64_t)M0 * c.v[0];
    const uint64_t p1 = (uint64_t)M1 * c.v[2];
    const uint32_t hi0 = (uint32_t)(p0 >> 32), lo0 = (uint32_t)p0;
    const uint32_t hi1 = (uint32_t)(p1 >> 32), lo1 = (uint32_t)p1;
    U4 n;
    n.v[0] = hi1 ^ c.v[1] ^ k0;
    n.v[1] = lo1;
    n.v[2] = hi0 ^ c.v[3] ^ k1;
    n.v[3] = lo0;
    c = n;
    k0 += W0;
    k1 += W1;
  }
  return c;
}

// 53-bit uniform in (0, 1) from two 32-bit words.
__device__ __forceinline__ double u53(uint32_t hi, uint32_t lo) {
  const uint64_t m = ((uint64_t)hi << 21) | (uint64_t)(lo >> 11);
  return ((double)m + 0.5) * 1.1102230246251565404236316680908203125e-16;  // 2^-53
}

// White noise on the half spectrum.  Modes on the k_z = 0 and k_z = n2/2
// planes pair with (-i, -j); the pair member with the smaller (i n1 + j) is
// generated and the other is its conjugate, so the field is Hermitian.
// Self-conjugate modes are real with unit variance.
template <typename T>
__global__ void k_whitenoise(C<T>* out, int n0, int n1, int n2, int64_t pitch2, uint32_t key0,
                             uint32_t key1, int unitary) {
  const int nh = n2 / 2 + 1;
  const int64_t total = (int64_t)n0 * n1 * nh;
  for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < total;
       t += (int64_t)gridDim.x * blockDim.x) {
    const int k = (int)(t % nh);
    const int64_t r = t / nh;
    const int j = (int)(r % n1), i = (int)(r / n1);
    int ic = i, jc = j;
    bool conj = false, self = false;
    const bool plane = (k == 0) || ((n2 % 2 == 0) && (k == n2 / 2));
    if (plane) {
      const int ip = (n0 - i) % n0, jp = (n1 - j) % n1;
      const int64_t lin = (int64_t)i * n1 + j, linp = (int64_t)ip * n1 + jp;
      if (linp < lin) {
        ic = ip;
        jc = jp;
        conj = true;
      } else if (linp == lin) {
        self = true;
      }
    }
    const uint64_t ctr = ((uint64_t)ic * (uint64_t)n1 + (uint64_t)jc) * (uint64_t)nh + (uint64_t)k;
    U4 c;
    c.v[0] = (uint32_t)ctr;
    c.v[1] = (uint32_t)(ctr >> 32);
    c.v[2] = 0u;
    c.v[3] = 0u;
    const U4 o = philox4x32_10(c, key0, key1);
    const double ua = u53(o.v[0], o.v[1]);
    const double ub = u53(o.v[2], o.v[3]);
    double cs, sn;
    sincospi(2.0 * ub, &sn, &cs);
    double re, im;
    if (unitary) {
      re = cs;
      im = sn;
    } else {
      const double amp = sqrt(-log(ua));
      re = amp * cs;
      im = amp * sn;
    }
    if (self) {
      re = unitary ? (cs >= 0.0 ? 1.0 : -1.0) : 1.4142135623730950488016887242097 * re;
      im = 0.0;
    }
    if (conj) im = -im;
    if (i == 0 && j == 0 && k == 0) {
      re = 0.0;
      im = 0.0;
    }
    out[r * pitch2 + k] = cmk<C<T>>((T)re, (T)im);
  }
}

__device__ __forceinline__ int fftfreq_i(int i, int n) { return i < (n + 1) / 2 ? i : i - n; }

template <typename T>
__global__ void k_linear_scale(C<T>* f, int n0, int n1, int n2, int64_t pitch2, const double* tab,
                               int64_t ntab) {
  const int nh = n2 / 2 + 1;
  const int64_t total = (int64_t)n0 * n1 * nh;
  for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < total;
       t += (int64_t)gridDim.x * blockDim.x) {
    const int k = (int)(t % nh);
    const int64_t r = t / nh;
    const int j = (int)(r % n1), i = (int)(r / n1);
    const int64_t a = fftfreq_i(i, n0), b = fftfreq_i(j, n1);
    int64_t n2s = a * a + b * b + (int64_t)k * k;
    if (n2s >= ntab) n2s = ntab - 1;
    const double s = tab[n2s];
    C<T>* p = f + r * pitch2 + k;
    const C<T> v = *p;
    *p = cmk<C<T>>((T)((double)v.x * s), (T)((double)v.y * s));
  }
}

// The same scaling with the table entry of each mode named by an index array
// ([n0][n1][n2/2+1] int32, one entry per half-spectrum mode): boxes whose
// sides differ, where |k| takes too many values for an n^2 table.
template <typename T>
__global__ void k_linear_scale_indexed(C<T>* f, int n0, int n1, int n2, int64_t pitch2,
                                       const double* tab, const int32_t* index) {
  const int nh = n2 / 2 + 1;
  const int64_t total = (int64_t)n0 * n1 * nh;
  for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < total;
       t += (int64_t)gridDim.x * blockDim.x) {
    const int k = (int)(t % nh);
    const int64_t r = t / nh;
    const double s = tab[index[t]];
    C<T>* p = f + r * pitch2 + k;
    const C<T> v = *p;
    *p = cmk<C<T>>((T)((double)v.x * s), (T)((double)v.y * s));
  }
}

template <typename T>
__global__ void k_lpt2_source(const T* xx, const T* yy, const T* zz, const T* xy, const T* yz,
                              const T* zx, T* out, int64_t n) {
  const T c = (T)(3.0 / 7.0);
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x) {
    const T a = xx[i], b = yy[i], d = zz[i];
    const T e = xy[i], f = yz[i], g = zx[i];
    out[i] = c * (a * b + b * d + d * a - e * e - f * f - g * g);
  }
}

// ---------------------------------------------------------------- binned spectra
// Each 256-thread block walks a contiguous range of modes.  Within a wave the
// lanes that share a bin are summed with a fixed xor-butterfly, and lane 0 adds
// the sum to that wave's LDS bins, so the order of every addition is fixed.
// Block partials go to `work`; a second kernel sums them in block order.
constexpr int kPowerThreads = 256;
constexpr int kPowerSums = 5;

__device__ __forceinline__ double wave_sum(double x) {
#pragma unroll
  for (int o = 32; o >= 1; o >>= 1) x += __shfl_xor(x, o);
  return x;
}

template <typename T>
__global__ __launch_bounds__(kPowerThreads) void k_power_partial(const C<T>* f1, const C<T>* f2,
                                                                  int n0, int n1, int n2,
                                                                  int64_t pitch2, double L0,
                                                                  double L1, double L2, int nbins,
                                                                  double kmin, double dk,
                                                                  int64_t chunk, double* work) {
  extern __shared__ double acc[];  // [4 waves][nbins][5]
  const int wave = threadIdx.x / 64, lane = threadIdx.x % 64;
  const int nw = kPowerThreads / 64;
  for (int i = threadIdx.x; i < nw * nbins * kPowerSums; i += kPowerThreads) acc[i] = 0.0;
  __syncthreads();
  double* my = acc + wave * nbins * kPowerSums;
  const int nh = n2 / 2 + 1;
  const int64_t total = (int64_t)n0 * n1 * nh;
  const int64_t start = (int64_t)blockIdx.x * chunk;
  int64_t end = start + chunk;
  if (end > total) end = total;
  const double two_pi = 6.283185307179586476925286766559;
  for (int64_t base = start; base < end; base += kPowerThreads) {
    const int64_t t = base + threadIdx.x;
    int bin = -1;
    double v[kPowerSums] = {0, 0, 0, 0, 0};
    if (t < end) {
      const int k = (int)(t % nh);
      const int64_t r = t / nh;
      const int j = (int)(r % n1), i = (int)(r / n1);
      const double kx = two_pi * fftfreq_i(i, n0) / L0;
      const double ky = two_pi * fftfreq_i(j, n1) / L1;
      const double kz = two_pi * k / L2;
      const double kk = sqrt(kx * kx + ky * ky + kz * kz);
      const double fb = floor((kk - kmin) / dk);
      if (!(i == 0 && j == 0 && k == 0) && fb >= 0.0 && fb < (double)nbins) {
        bin = (int)fb;
        const double w = (k == 0 || ((n2 % 2 == 0) && k == n2 / 2)) ? 1.0 : 2.0;
        const C<T> a = f1[r * pitch2 + k];
        const C<T> b = f2 ? f2[r * pitch2 + k] : a;
        const double ax = a.x, ay = a.y, bx = b.x, by = b.y;
        v[0] = w;
        v[1] = w * kk;
        v[2] = w * (ax * bx + ay * by);
        v[3] = w * (ax * ax + ay * ay);
        v[4] = w * (bx * bx + by * by);
      }
    }
    unsigned long long active = __ballot(bin >= 0);
    while (active) {
      const int leader = __ffsll(active) - 1;
      const int b = __shfl(bin, leader);
      const bool mine = (bin == b);
#pragma unroll
      for (int s = 0; s < kPowerSums; ++s) {
        const double x = wave_sum(mine ? v[s] : 0.0);
        if (lane == 0) my[b * kPowerSums + s] += x;
      }
      active &= ~__ballot(mine);
    }
  }
  __syncthreads();
  for (int idx = threadIdx.x; idx < nbins * kPowerSums; idx += kPowerThreads) {
    double s = 0.0;
    for (int w = 0; w < nw; ++w) s += acc[w * nbins * kPowerSums + idx];
    work[(int64_t)blockIdx.x * nbins * kPowerSums + idx] = s;
  }
}

__global__ void k_power_final(const double* work, int nblk, int n, double* out) {
  for (int idx = blockIdx.x * blockDim.x + threadIdx.x; idx < n; idx += gridDim.x * blockDim.x) {
    double s = 0.0;
    for (int b = 0; b < nblk; ++b) s += work[(int64_t)b * n + idx];
    out[idx] = s;
  }
}

static int power_blocks(const fpm_mesh* m) {
  const int64_t total = (int64_t)m->nmesh[0] * m->nmesh[1] * (m->nmesh[2] / 2 + 1);
  int64_t nb = total / 8192;
  if (nb < 1) nb = 1;
  if (nb > 2048) nb = 2048;
  return (int)nb;
}

static int check_cmesh(const fpm_mesh* m, int64_t pitch2) {
  if (!m) return FPM_ERR_ARG;
  for (int d = 0; d < 3; ++d)
    if (m->nmesh[d] <= 0 || !(m->box[d] > 0)) return FPM_ERR_ARG;
  if (pitch2 < m->nmesh[2] / 2 + 1) return FPM_ERR_ARG;
  return FPM_OK;
}

template <typename T>
static int power_impl(const void* f1, const void* f2, const fpm_mesh* m, int64_t pitch2, int nbins,
                      double kmin, double dk, double* work, int64_t work_len, double* out,
                      hipStream_t st) {
  int rc = check_cmesh(m, pitch2);
  if (rc) return rc;
  if (!f1 || !work || !out || nbins < 1 || nbins > 1024 || !(dk > 0)) return FPM_ERR_ARG;
  const int nblk = power_blocks(m);
  if (work_len < (int64_t)nblk * nbins * kPowerSums) return FPM_ERR_WORKSPACE;
  const int64_t total = (int64_t)m->nmesh[0] * m->nmesh[1] * (m->nmesh[2] / 2 + 1);
  int64_t chunk = (total + nblk - 1) / nblk;
  chunk = ((chunk + kPowerThreads - 1) / kPowerThreads) * kPowerThreads;
  const size_t lds = (size_t)(kPowerThreads / 64) * nbins * kPowerSums * sizeof(double);
  if (lds > 160 * 1024) return FPM_ERR_ARG;
  hipLaunchKernelGGL(k_power_partial<T>, dim3(nblk), dim3(kPowerThreads), lds, st,
                     static_cast<const C<T>*>(f1), static_cast<const C<T>*>(f2), m->nmesh[0],
                     m->nmesh[1], m->nmesh[2], pitch2, m->box[0], m->box[1], m->box[2], nbins, kmin,
                     dk, chunk, work);
  FPM_LAUNCH_CHECK();
  const int n = nbins * kPowerSums;
  hipLaunchKernelGGL(k_power_final, dim3((n + 255) / 256), dim3(256), 0, st, work, nblk, n, out);
  FPM_LAUNCH_CHECK();
  return FPM_OK;
}

// ---------------------------------------------------------------- allocation-free P(k)
// fpm_power_bin_noalloc keeps the block partials in this array instead of a
// caller workspace: 2048 blocks x 128 bins x 5 sums (10 MiB), held by every
// device the library runs on until the process ends.  Host code never passes
// the array's name to a kernel (on the host it is only a stand-in for the
// symbol): each device's address comes from hipGetSymbolAddress, queried once
// per device with that device current, on the entry point's first call there.
constexpr int64_t kPowerScratchDoubles = (int64_t)2048 * 128 * kPowerSums;
static __device__ __align__(256) double scratch[kPowerScratchDoubles];

constexpr int kMaxDevices = 64;
static std::mutex g_scratch_mu;
static double* g_scratch_dev[kMaxDevices] = {};

static int scratch_address(double** out) {
  int dev = -1;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= kMaxDevices) return FPM_ERR_HIP;
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_dev[dev]) {
    void* p = nullptr;
    if (hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)) != hipSuccess || !p) return FPM_ERR_HIP;
    g_scratch_dev[dev] = static_cast<double*>(p);
  }
  *out = g_scratch_dev[dev];
  return FPM_OK;
}

// The same kernels as power_impl, with the partials in `scratch`.  Every
// partial the final sum reads is written by the partial pass of the same
// call, so nothing needs clearing; calls that overlap on one device share the
// array and race.
template <typename T>
static int power_noalloc_impl(const void* f1, const void* f2, const fpm_mesh* m, int64_t pitch2,
                              int nbins, double kmin, double dk, double* out, hipStream_t st) {
  int rc = check_cmesh(m, pitch2);
  if (rc) return rc;
  if (nbins < 1) return FPM_ERR_ARG;
  if ((int64_t)power_blocks(m) * nbins * kPowerSums > kPowerScratchDoubles) return FPM_ERR_SCRATCH;
  double* work = nullptr;
  rc = scratch_address(&work);
  if (rc) return rc;
  return power_impl<T>(f1, f2, m, pitch2, nbins, kmin, dk, work, kPowerScratchDoubles, out, st);
}

template <typename T>
static int whitenoise_impl(void* out, const fpm_mesh* m, int64_t pitch2, uint64_t seed,
                           int unitary, hipStream_t st) {
  int rc = check_cmesh(m, pitch2);
  if (rc) return rc;
  if (!out) return FPM_ERR_ARG;
  const int64_t total = (int64_t)m->nmesh[0] * m->nmesh[1] * (m->nmesh[2] / 2 + 1);
  hipLaunchKernelGGL(k_whitenoise<T>, dim3(grid_elems(total)), dim3(256), 0, st,
                     static_cast<C<T>*>(out), m->nmesh[0], m->nmesh[1], m->nmesh[2], pitch2,
                     (uint32_t)seed, (uint32_t)(seed >> 32), unitary ? 1 : 0);
  FPM_LAUNCH_CHECK();
  return FPM_OK;
}

template <typename T>
static int linear_impl(void* f, const fpm_mesh* m, int64_t pitch2, const double* tab, int64_t ntab,
                       hipStream_t st) {
  int rc = check_cmesh(m, pitch2);
  if (rc) return rc;
  if (!f || !tab) return FPM_ERR_ARG;
  int64_t need = 0;
  for (int d = 0; d < 2; ++d) need += (int64_t)(m->nmesh[d] / 2) * (m->nmesh[d] / 2);
  need += (int64_t)(m->nmesh[2] / 2) * (m->nmesh[2] / 2);
  if (ntab < need + 1) return FPM_ERR_ARG;
  const int64_t total = (int64_t)m->nmesh[0] * m->nmesh[1] * (m->nmesh[2] / 2 + 1);
  hipLaunchKernelGGL(k_linear_scale<T>, dim3(grid_elems(total)), dim3(256), 0, st,
                     static_cast<C<T>*>(f), m->nmesh[0], m->nmesh[1], m->nmesh[2], pitch2, tab,
                     ntab);
  FPM_LAUNCH_CHECK();
  return FPM_OK;
}

template <typename T>
static int linear_indexed_impl(void* f, const fpm_mesh* m, int64_t pitch2, const double* tab,
                               int64_t ntab, const int32_t* index, hipStream_t st) {
  int rc = check_cmesh(m, pitch2);
  if (rc) return rc;
  if (!f || !tab || !index || ntab < 1) return FPM_ERR_ARG;
  const int n2 = m->nmesh[2];
  const int64_t total = (int64_t)m->nmesh[0] * m->nmesh[1] * (n2 / 2 + 1);
  hipLaunchKernelGGL(k_linear_scale_indexed<T>, dim3(grid_elems(total)), dim3(256), 0, st,
                     static_cast<C<T>*>(f), m->nmesh[0], m->nmesh[1], n2, pitch2, tab, index);
  FPM_LAUNCH_CHECK();
  return FPM_OK;
}

template <typename T>
static int lpt2_impl(const void* const* phi, void* out, int64_t n, hipStream_t st) {
  if (!phi || !out || n < 0) return FPM_ERR_ARG;
  for (int i = 0; i < 6; ++i)
    if (!phi[i]) return FPM_ERR_ARG;
  if (n == 0) return FPM_OK;
  hipLaunchKernelGGL(k_lpt2_source<T>, dim3(grid_elems(n)), dim3(256), 0, st,
                     static_cast<const T*>(phi[0]), static_cast<const T*>(phi[1]),
                     static_cast<const T*>(phi[2]), static_cast<const T*>(phi[3]),
                     static_cast<const T*>(phi[4]), static_cast<const T*>(phi[5]),
                     static_cast<T*>(out), n);
  FPM_LAUNCH_CHECK();
  return FPM_OK;
}

}  // namespace fpm

using namespace fpm;

extern "C" {

int fpm_whitenoise_f32(void* out, const fpm_mesh* m, int64_t p, uint64_t seed, int u,
                       fpm_stream_t s) {
  return whitenoise_impl<float>(out, m, p, seed, u, (hipStream_t)s);
}
int fpm_whitenoise_f64(void* out, const fpm_mesh* m, int64_t p, uint64_t seed, int u,
                       fpm_stream_t s) {
  return whitenoise_impl<double>(out, m, p, seed, u, (hipStream_t)s);
}
int fpm_linear_scale_f32(void* f, const fpm_mesh* m, int64_t p, const double* t, int64_t nt,
                         fpm_stream_t s) {
  return linear_impl<float>(f, m, p, t, nt, (hipStream_t)s);
}
int fpm_linear_scale_f64(void* f, const fpm_mesh* m, int64_t p, const double* t, int64_t nt,
                         fpm_stream_t s) {
  return linear_impl<double>(f, m, p, t, nt, (hipStream_t)s);
}
int fpm_linear_scale_indexed_f32(void* f, const fpm_mesh* m, int64_t p, const double* t,
                                 int64_t nt, const int32_t* idx, fpm_stream_t s) {
  return linear_indexed_impl<float>(f, m, p, t, nt, idx, (hipStream_t)s);
}
int fpm_linear_scale_indexed_f64(void* f, const fpm_mesh* m, int64_t p, const double* t,
                                 int64_t nt, const int32_t* idx, fpm_stream_t s) {
  return linear_indexed_impl<double>(f, m, p, t, nt, idx, (hipStream_t)s);
}
int fpm_lpt2_source_f32(const void* const* phi, void* out, int64_t n, fpm_stream_t s) {
  return lpt2_impl<float>(phi, out, n, (hipStream_t)s);
}
int fpm_lpt2_source_f64(const void* const* phi, void* out, int64_t n, fpm_stream_t s) {
  return lpt2_impl<double>(phi, out, n, (hipStream_t)s);
}
int64_t fpm_power_work_len(const fpm_mesh* m, int nbins) {
  if (!m || nbins < 1) return 0;
  return (int64_t)power_blocks(m) * nbins * kPowerSums;
}
int fpm_power_bin_f32(const void* f1, const void* f2, const fpm_mesh* m, int64_t p, int nbins,
                      double kmin, double dk, double* w, int64_t wl, double* out, fpm_stream_t s) {
  return power_impl<float>(f1, f2, m, p, nbins, kmin, dk, w, wl, out, (hipStream_t)s);
}
int fpm_power_bin_f64(const void* f1, const void* f2, const fpm_mesh* m, int64_t p, int nbins,
                      double kmin, double dk, double* w, int64_t wl, double* out, fpm_stream_t s) {
  return power_impl<double>(f1, f2, m, p, nbins, kmin, dk, w, wl, out, (hipStream_t)s);
}
int fpm_power_bin_noalloc_f32(const void* f1, const void* f2, const fpm_mesh* m, int64_t p,
                              int nbins, double kmin, double dk, double* out, fpm_stream_t s) {
  return power_noalloc_impl<float>(f1, f2, m, p, nbins, kmin, dk, out, (hipStream_t)s);
}
int fpm_power_bin_noalloc_f64(const void* f1, const void* f2, const fpm_mesh* m, int64_t p,
                              int nbins, double kmin, double dk, double* out, fpm_stream_t s) {
  return power_noalloc_impl<double>(f1, f2, m, p, nbins, kmin, dk, out, (hipStream_t)s);
}

}  // extern "C"
