// Host-side growth functions for kick / drift coefficients (double precision).
//
// Solves, in ln a, the first- and second-order growth equations of a matter +
// Lambda (+ curvature) universe (fastpm/background.py:175-184):
//   D1'' = h D1' + 3/2 Om(a) D1
//   D2'' = h D2' + 3/2 Om(a) (D2 - D1^2),      h = -2 - E'/E,
// from the matter-dominated start (a0, a0, -3/7 a0^2, -6/7 a0^2) at
// a0 = min(1e-7, a) (fastpm/background.py:243-247), then scales D1 and D2
// (with their derivatives) to 1 at a_normalize (fastpm/background.py:201-204).
// Integration: classical fourth-order Runge-Kutta with steps of at most 1e-3 in
// ln a, landing exactly on every requested a.
#include <algorithm>
#include <cmath>
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

void rk4(const Cosmo& c, double x0, double x1, double y[4]) {
  const double span = x1 - x0;
  if (span == 0.0) return;
  const int n = std::max(1, (int)std::ceil(std::fabs(span) / 1e-3));
  const double h = span / n;
  double k1[4], k2[4], k3[4], k4[4], t[4];
  double x = x0;
  for (int s = 0; s < n; ++s) {
    c.rhs(x, y, k1);
    for (int i = 0; i < 4; ++i) t[i] = y[i] + 0.5 * h * k1[i];
    c.rhs(x + 0.5 * h, t, k2);
    for (int i = 0; i < 4; ++i) t[i] = y[i] + 0.5 * h * k2[i];
    c.rhs(x + 0.5 * h, t, k3);
    for (int i = 0; i < 4; ++i) t[i] = y[i] + h * k3[i];
    c.rhs(x + h, t, k4);
    for (int i = 0; i < 4; ++i) y[i] += h / 6.0 * (k1[i] + 2.0 * k2[i] + 2.0 * k3[i] + k4[i]);
    x = x0 + (s + 1) * h;
  }
}

}  // namespace

extern "C" int fpm_growth(double omega_m, double omega_lambda, double omega_k, double a_normalize,
                          const double* a, int n, double* out) {
  if (n < 0 || (n > 0 && (!a || !out)) || !(a_normalize > 0)) return FPM_ERR_ARG;
  for (int i = 0; i < n; ++i)
    if (!(a[i] > 0)) return FPM_ERR_ARG;
  const Cosmo c{omega_m, omega_lambda, omega_k};
  // points to visit: a_normalize and every request, in increasing order
  std::vector<double> pts(a, a + n);
  pts.push_back(a_normalize);
  std::sort(pts.begin(), pts.end());
  pts.erase(std::unique(pts.begin(), pts.end()), pts.end());
  const double a0 = std::min(1e-7, pts.front());
  double y[4] = {a0, a0, -3.0 / 7.0 * a0 * a0, -6.0 / 7.0 * a0 * a0};
  double x = std::log(a0);
  std::vector<double> sol(pts.size() * 6);
  for (size_t p = 0; p < pts.size(); ++p) {
    const double xt = std::log(pts[p]);
    rk4(c, x, xt, y);
    x = xt;
    double dy[4];
    c.rhs(xt, y, dy);
    double* s = &sol[p * 6];
    s[0] = y[0];
    s[1] = y[1];
    s[2] = dy[1];
    s[3] = y[2];
    s[4] = y[3];
    s[5] = dy[3];
  }
  const size_t inorm = std::lower_bound(pts.begin(), pts.end(), a_normalize) - pts.begin();
  const double d1n = sol[inorm * 6 + 0], d2n = sol[inorm * 6 + 3];
  for (int i = 0; i < n; ++i) {
    const size_t p = std::lower_bound(pts.begin(), pts.end(), a[i]) - pts.begin();
    const double* s = &sol[p * 6];
    double* o = out + 9 * (size_t)i;
    o[0] = s[0] / d1n;
    o[1] = s[1] / d1n;
    o[2] = s[2] / d1n;
    o[3] = s[3] / d2n;
    o[4] = s[4] / d2n;
    o[5] = s[5] / d2n;
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
    case FPM_ERR_UNSUPPORTED: return "unsupported FFT length";
    case FPM_ERR_HIP: return "HIP launch failed";
    case FPM_ERR_WORKSPACE: return "workspace too small";
    default: return "unknown status";
  }
}
