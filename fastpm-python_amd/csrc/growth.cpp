// Host-side growth functions for kick / drift coefficients (double precision).
//
// Solves, in ln a, the first- and second-order growth equations of a matter +
// Lambda (+ curvature) universe (fastpm/background.py:175-184):
//   D1'' = h D1' + 3/2 Om(a) D1
//   D2'' = h D2' + 3/2 Om(a) (D2 - D1^2),      h = -2 - E'/E,
// from the matter-dominated start (a0, a0, -3/7 a0^2, -6/7 a0^2) at
// a0 = min(1e-7, a) (fastpm/background.py:243-247), then scales D1 and D2
// (with their derivatives) to 1 at a_normalize (fastpm/background.py:201-204).
// Integration: Dormand-Prince 5 steps of fixed size 1/256 in ln a from
// ln 1e-7, kept per cosmology, and one partial step from the grid point
// below each requested a, so the value at a depends on a alone (not on which
// other scale factors are requested with it) and a query costs one step once
// the grid reaches a.
#include <algorithm>
#include <cmath>
#include <mutex>
#include <vector>

#include "../../include/fpm.h"

namespace {

struct Cosmo {
  double om, ol, ok;
  double efunc(double a) const { return std::sqrt(om / (a * a * a) + ok / (a * a) + ol); }
  // dE/d ln a
  double eprime(double a) const {
    return a * 0.5 / efunc(a) * (-3.0 * om / (a * a * a * a) - 2.0 * ok / (a * a * a));
  }
  double omega_m(double a) const {
    const double e = efunc(a);
    return om / (a * a * a) / (e * e);
  }
  void rhs(double lna, const double y[4], double dy[4]) const {
    const double a = std::exp(lna);
    const double h = -2.0 - eprime(a) / efunc(a);
    const double o = 1.5 * omega_m(a);
    dy[0] = y[1];
    dy[1] = h * y[1] + o * y[0];
    dy[2] = y[3];
    dy[3] = h * y[3] + o * y[2] - o * y[0] * y[0];
  }
};

// One Dormand-Prince 5 step (Dormand & Prince 1980) of size h from (x, y),
// without error control.
void dp_step(const Cosmo& c, double x, double h, const double y[4], double yn[4]) {
  static const double c2 = 1.0 / 5, c3 = 3.0 / 10, c4 = 4.0 / 5, c5 = 8.0 / 9;
  static const double a21 = 1.0 / 5;
  static const double a31 = 3.0 / 40, a32 = 9.0 / 40;
  static const double a41 = 44.0 / 45, a42 = -56.0 / 15, a43 = 32.0 / 9;
  static const double a51 = 19372.0 / 6561, a52 = -25360.0 / 2187, a53 = 64448.0 / 6561,
                      a54 = -212.0 / 729;
  static const double a61 = 9017.0 / 3168, a62 = -355.0 / 33, a63 = 46732.0 / 5247,
                      a64 = 49.0 / 176, a65 = -5103.0 / 18656;
  static const double b1 = 35.0 / 384, b3 = 500.0 / 1113, b4 = 125.0 / 192,
                      b5 = -2187.0 / 6784, b6 = 11.0 / 84;
  double k1[4], k2[4], k3[4], k4[4], k5[4], k6[4], t[4];
  c.rhs(x, y, k1);
  for (int i = 0; i < 4; ++i) t[i] = y[i] + h * a21 * k1[i];
  c.rhs(x + c2 * h, t, k2);
  for (int i = 0; i < 4; ++i) t[i] = y[i] + h * (a31 * k1[i] + a32 * k2[i]);
  c.rhs(x + c3 * h, t, k3);
  for (int i = 0; i < 4; ++i) t[i] = y[i] + h * (a41 * k1[i] + a42 * k2[i] + a43 * k3[i]);
  c.rhs(x + c4 * h, t, k4);
  for (int i = 0; i < 4; ++i)
    t[i] = y[i] + h * (a51 * k1[i] + a52 * k2[i] + a53 * k3[i] + a54 * k4[i]);
  c.rhs(x + c5 * h, t, k5);
  for (int i = 0; i < 4; ++i)
    t[i] = y[i] + h * (a61 * k1[i] + a62 * k2[i] + a63 * k3[i] + a64 * k4[i] + a65 * k5[i]);
  c.rhs(x + h, t, k6);
  for (int i = 0; i < 4; ++i)
    yn[i] = y[i] + h * (b1 * k1[i] + b3 * k3[i] + b4 * k4[i] + b5 * k5[i] + b6 * k6[i]);
}

// The solution on the fixed grid x_k = X0 + k H, one track per cosmology,
// extended on demand (guarded: Python may call from several threads).
constexpr double kA0 = 1e-7;
constexpr double kH = 1.0 / 256;
struct Track {
  double om, ol, ok;
  std::vector<double> y;  // 4 per grid point
};
std::mutex g_mu;
std::vector<Track> g_tracks;

// State at ln a = x >= X0 (the caller holds g_mu).
void state_at(Track& t, const Cosmo& c, double x, double out[4]) {
  const double x0 = std::log(kA0);
  if (t.y.empty()) {
    t.y = {kA0, kA0, -3.0 / 7.0 * kA0 * kA0, -6.0 / 7.0 * kA0 * kA0};
  }
  const int64_t k = (int64_t)std::floor((x - x0) / kH);
  while ((int64_t)(t.y.size() / 4) <= k) {
    const int64_t j = (int64_t)(t.y.size() / 4) - 1;
    double yn[4];
    dp_step(c, x0 + (double)j * kH, kH, &t.y[4 * j], yn);
    t.y.insert(t.y.end(), yn, yn + 4);
  }
  const double xk = x0 + (double)k * kH;
  if (x > xk) {
    dp_step(c, xk, x - xk, &t.y[4 * k], out);
  } else {
    for (int i = 0; i < 4; ++i) out[i] = t.y[4 * k + i];
  }
}

// D1, D1', D2, D2' (unnormalised) at a: the matter-dominated start below a0.
void growth_at(Track& t, const Cosmo& c, double a, double y[4]) {
  if (a <= kA0) {
    y[0] = a;
    y[1] = a;
    y[2] = -3.0 / 7.0 * a * a;
    y[3] = -6.0 / 7.0 * a * a;
    return;
  }
  state_at(t, c, std::log(a), y);
}

}  // namespace

extern "C" int fpm_growth(double omega_m, double omega_lambda, double omega_k, double a_normalize,
                          const double* a, int n, double* out) {
  if (n < 0 || (n > 0 && (!a || !out)) || !(a_normalize > 0)) return FPM_ERR_ARG;
  for (int i = 0; i < n; ++i)
    if (!(a[i] > 0)) return FPM_ERR_ARG;
  const Cosmo c{omega_m, omega_lambda, omega_k};
  std::lock_guard<std::mutex> lock(g_mu);
  Track* t = nullptr;
  for (auto& tr : g_tracks)
    if (tr.om == omega_m && tr.ol == omega_lambda && tr.ok == omega_k) t = &tr;
  if (!t) {
    if (g_tracks.size() >= 64) g_tracks.clear();  // bounded: re-solved on demand
    g_tracks.push_back(Track{omega_m, omega_lambda, omega_k, {}});
    t = &g_tracks.back();
  }
  double yn[4];
  growth_at(*t, c, a_normalize, yn);
  const double d1n = yn[0], d2n = yn[2];
  for (int i = 0; i < n; ++i) {
    double y[4], dy[4];
    growth_at(*t, c, a[i], y);
    c.rhs(std::log(a[i]), y, dy);
    double* o = out + 9 * (size_t)i;
    o[0] = y[0] / d1n;
    o[1] = y[1] / d1n;
    o[2] = dy[1] / d1n;
    o[3] = y[2] / d2n;
    o[4] = y[3] / d2n;
    o[5] = dy[3] / d2n;
    o[6] = c.efunc(a[i]);
    o[7] = c.eprime(a[i]);
    o[8] = c.omega_m(a[i]);
  }
  return FPM_OK;
}

extern "C" const char* fpm_version(void) { return "fpm-mi355x 0.1"; }

extern "C" const char* fpm_strerror(int status) {
  switch (status) {
    case FPM_OK: return "ok";
    case FPM_ERR_ARG: return "invalid argument (shape, size, order or pointer)";
    case FPM_ERR_UNSUPPORTED: return "unsupported size (FFT length or mesh side)";
    case FPM_ERR_HIP: return "HIP launch failed";
    case FPM_ERR_WORKSPACE: return "workspace too small";
    case FPM_ERR_SCRATCH: return "request exceeds the fixed scratch of an allocation-free call";
    default: return "unknown status";
  }
}
