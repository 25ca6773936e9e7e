"""Growth functions and FastPM kick/drift factors (reference:
fastpm/background.py:4-256).

The ODE is solved by the native library (fpm_growth: Dormand-Prince 5 steps
of 1/256 in ln a from a = 1e-7, kept per cosmology, and one partial step to
each queried scale factor; double precision), so values at the grid points
of the reference agree with its odeint solution to the ODE tolerance, values
off the grid are solved rather than interpolated, and a value does not
depend on which other scale factors are queried with it.
"""
import numpy as np

from . import _lib


class MatterDominated:
    """Single-fluid growth in a matter + Lambda (+ curvature) universe.

    D1 and D2 are each normalised to 1 at ``a_normalize``; derivatives are with
    respect to ln a (background.py:39-173)."""

    def __init__(self, Omega0_m, Omega0_lambda=None, Omega0_k=0, a=None, a_normalize=1.0):
        if Omega0_lambda is None:
            Omega0_lambda = 1 - Omega0_k - Omega0_m
        self.Omega0_m = float(Omega0_m)
        self.Omega0_lambda = float(Omega0_lambda)
        self.Omega0_k = float(Omega0_k)
        self.Om0 = self.Omega0_m
        self.a_normalize = float(a_normalize)
        self._cache = {}
        if a is not None:
            self._rows(np.atleast_1d(np.asarray(a, dtype="f8")))

    def _rows(self, a):
        a = np.atleast_1d(np.asarray(a, dtype="f8"))
        missing = sorted({float(x) for x in a.ravel() if float(x) not in self._cache})
        if missing:
            tab = _lib.growth(self.Omega0_m, self.Omega0_lambda, self.Omega0_k,
                              self.a_normalize, missing)
            for x, row in zip(missing, tab):
                self._cache[x] = row
        return np.array([self._cache[float(x)] for x in a.ravel()]).reshape(a.shape + (9,))

    def _col(self, a, i):
        scalar = np.ndim(a) == 0
        v = self._rows(a)[..., i]
        return float(v.ravel()[0]) if scalar else v

    def D1(self, a, order=0):
        return self._col(a, order)

    def D2(self, a, order=0):
        return self._col(a, 3 + order)

    def E(self, a, order=0):
        return self._col(a, 6 if order == 0 else 7)

    def Om(self, a):
        return self._col(a, 8)

    def efunc(self, a):
        return self.E(a)

    def f1(self, a):
        return self.D1(a, 1) / self.D1(a)

    def f2(self, a):
        return self.D2(a, 1) / self.D2(a)

    def Gp(self, a):
        return self.D1(a)

    def Gp2(self, a):
        return self.D2(a)

    def gp(self, a):
        return self.D1(a, 1) / a

    def gp2(self, a):
        return self.D2(a, 1) / a

    def Gf(self, a):
        return self.D1(a, 1) * a ** 2 * self.E(a)

    def Gf2(self, a):
        return self.D2(a, 1) * a ** 2 * self.E(a)

    def gf(self, a):
        return 1 / a * (self.D1(a, 2) * a ** 2 * self.E(a)
                        + self.D1(a, 1) * (a ** 2 * self.E(a, 1) + 2 * a ** 2 * self.E(a)))

    def gf2(self, a):
        return 1 / a * (self.D2(a, 2) * a ** 2 * self.E(a)
                        + self.D2(a, 1) * (a ** 2 * self.E(a, 1) + 2 * a ** 2 * self.E(a)))

    # FastPM stepping factors (fastpm/core.py:168-180)
    def kick_factor(self, ai, ac, af):
        return 1 / (ac ** 2 * self.E(ac)) * (self.Gf(af) - self.Gf(ai)) / self.gf(ac)

    def drift_factor(self, ai, ac, af):
        return 1 / (ac ** 3 * self.E(ac)) * (self.Gp(af) - self.Gp(ai)) / self.gp(ac)


PerturbationGrowth = MatterDominated
