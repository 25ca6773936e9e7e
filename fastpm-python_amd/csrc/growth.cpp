// Host-side growth functions for kick / drift coefficients (double precision).
//
// Solves, in ln a, the first- and second-order growth equations of a matter +
// Lambda (+ curvature) universe (fastpm/background.py:175-184):
//   D1'' = h D1' + 3/2 Om(a) D1
//   D2'' = h D2' + 3/2 Om(a) (D2 - D1^2),      h = -2 - E'/E,
// from the matter-dominated start (a0, a0, -3/7 a0^2, -6/7 a0^2) at
// a0 = min(1e-7, a) (fastpm/background.py:243-247), then scales D1 and D2
// (with their derivatives) to 1 at a_normalize (fastpm/background.py:201-204).
// Integration: adaptive Dormand-Prince 5(4), landing exactly on every
// requested a.
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

// Dormand-Prince 5(4) embedded Runge-Kutta (Dormand & Prince 1980) with
// step-size control on a mixed absolute/relative error of 1e-13, landing
// exactly on x1.
void integrate(const Cosmo& c, double x0, double x1, double y[4]) {
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
  static const double e1 = 35.0 / 384 - 5179.0 / 57600, e3 = 500.0 / 1113 - 7571.0 / 16695,
                      e4 = 125.0 / 192 - 393.0 / 640, e5 = -2187.0 / 6784 + 92097.0 / 339200,
                      e6 = 11.0 / 84 - 187.0 / 2100, e7 = -1.0 / 40;
  const double tol = 1e-13;
  double x = x0;
  double h = std::min(0.01, x1 - x0);
  double k1[4], k2[4], k3[4], k4[4], k5[4], k6[4], k7[4], t[4], yn[4];
  int guard = 0;
  while (x < x1 && guard++ < 1000000) {
    if (x + h > x1) h = x1 - x;
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
    c.rhs(x + h, yn, k7);
    double err = 0.0;
    for (int i = 0; i < 4; ++i) {
      const double ei = h * (e1 * k1[i] + e3 * k3[i] + e4 * k4[i] + e5 * k5[i] + e6 * k6[i] +
                             e7 * k7[i]);
      const double sc = tol * (1e-30 + std::max(std::fabs(y[i]), std::fabs(yn[i])));
      err = std::max(err, std::fabs(ei) / sc);
    }
    if (err <= 1.0) {
      x += h;
      for (int i = 0; i < 4; ++i) y[i] = yn[i];
    }
    const double fac = err > 0 ? 0.9 * std::pow(err, -0.2) : 5.0;
    h *= std::min(5.0, std::max(0.2, fac));
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
    integrate(c, x, xt, y);
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
    case FPM_ERR_UNSUPPORTED: return "unsupported size (FFT length or mesh side)";
    case FPM_ERR_HIP: return "HIP launch failed";
    case FPM_ERR_WORKSPACE: return "workspace too small";
    default: return "unknown status";
  }
}
