"""Background cosmology and linear power for driving the solver.

The reference takes an nbodykit ``Cosmology`` (only ``Om0`` enters the solver,
fastpm/core.py:99-100,127,170,177,207) and a host callable P(k).  nbodykit is
not available, so this module supplies a small flat-LCDM description and the
Eisenstein & Hu (1998) no-wiggle linear power spectrum, normalised to sigma8.
"""
import numpy as np


class Cosmology:
    """Flat (or curved) LCDM parameters; h = H0 / 100."""

    def __init__(self, h=0.7, Om0=0.3, Ob0=0.05, n_s=0.96, sigma8=0.8, Tcmb0=2.7255,
                 Neff=3.046, Ok0=0.0):
        self.h = float(h)
        self.H0 = 100.0 * self.h
        self.Om0 = float(Om0)
        self.Ob0 = float(Ob0)
        self.n_s = float(n_s)
        self.sigma8 = float(sigma8)
        self.Tcmb0 = float(Tcmb0)
        self.Neff = float(Neff)
        self.Ok0 = float(Ok0)
        self.Ode0 = 1.0 - self.Om0 - self.Ok0

    def efunc(self, z):
        a = 1.0 / (1.0 + np.asarray(z, dtype="f8"))
        return np.sqrt(self.Om0 / a ** 3 + self.Ok0 / a ** 2 + self.Ode0)

    def scale_independent_growth_factor(self, z):
        """D1(a) normalised to 1 today."""
        from .background import MatterDominated

        a = 1.0 / (1.0 + np.asarray(z, dtype="f8"))
        return MatterDominated(self.Om0, self.Ode0, self.Ok0, a_normalize=1.0).D1(a)

    def clone(self, **kw):
        args = dict(h=self.h, Om0=self.Om0, Ob0=self.Ob0, n_s=self.n_s, sigma8=self.sigma8,
                    Tcmb0=self.Tcmb0, Neff=self.Neff, Ok0=self.Ok0)
        args.update(kw)
        return Cosmology(**args)


# Planck 2015 (TT,TE,EE+lowP+lensing+ext) values, the cosmology the reference's
# tests and CLI default to.
Planck15 = Cosmology(h=0.6774, Om0=0.3075, Ob0=0.0486, n_s=0.9667, sigma8=0.8159,
                     Tcmb0=2.7255, Neff=3.046)


def eh_nowiggle_transfer(k, cosmo):
    """Eisenstein & Hu (1998, ApJ 496, 605) zero-baryon-oscillation transfer
    function, eqs. 26 and 28-31; k in h/Mpc."""
    k = np.asarray(k, dtype="f8")
    h = cosmo.h
    omh2 = cosmo.Om0 * h * h
    obh2 = cosmo.Ob0 * h * h
    fb = cosmo.Ob0 / cosmo.Om0
    theta = cosmo.Tcmb0 / 2.7
    s = 44.5 * np.log(9.83 / omh2) / np.sqrt(1 + 10 * obh2 ** 0.75)  # Mpc
    alpha = 1 - 0.328 * np.log(431 * omh2) * fb + 0.38 * np.log(22.3 * omh2) * fb ** 2
    gamma_eff = cosmo.Om0 * h * (alpha + (1 - alpha) / (1 + (0.43 * k * h * s) ** 4))
    with np.errstate(divide="ignore", invalid="ignore"):
        q = k * theta ** 2 / gamma_eff
        L0 = np.log(2 * np.e + 1.8 * q)
        C0 = 14.2 + 731.0 / (1 + 62.5 * q)
        T = L0 / (L0 + C0 * q * q)
    return np.where(k > 0, T, 1.0)


def _sigma_r(pk, r):
    from scipy.integrate import quad

    def integrand(lnk):
        k = np.exp(lnk)
        x = k * r
        w = 3 * (np.sin(x) - x * np.cos(x)) / x ** 3
        return k ** 3 * pk(k) * w * w / (2 * np.pi ** 2)

    v, _ = quad(integrand, np.log(1e-6), np.log(1e3), limit=1000, epsabs=0, epsrel=1e-10)
    return np.sqrt(v)


class LinearPower:
    """Linear matter power P(k) [(Mpc/h)^3] at redshift z, k in h/Mpc.

    transfer='NoWiggleEisensteinHu' (the only one available offline);
    normalised so that sigma_8(z=0) = cosmo.sigma8."""

    def __init__(self, cosmo, redshift=0.0, transfer="NoWiggleEisensteinHu"):
        if transfer not in ("NoWiggleEisensteinHu", "EHNoWiggle"):
            raise NotImplementedError("only the no-wiggle Eisenstein-Hu transfer is available")
        self.cosmo = cosmo
        self.redshift = float(redshift)
        base = lambda k: np.asarray(k, dtype="f8") ** cosmo.n_s * eh_nowiggle_transfer(k, cosmo) ** 2  # noqa: E731
        self._norm = (cosmo.sigma8 / _sigma_r(base, 8.0)) ** 2
        self._D = float(cosmo.scale_independent_growth_factor(self.redshift)) if redshift else 1.0
        self._base = base

    def __call__(self, k):
        k = np.asarray(k, dtype="f8")
        with np.errstate(invalid="ignore"):
            return np.where(k > 0, self._norm * self._D ** 2 * self._base(np.where(k > 0, k, 1.0)), 0.0)

    def sigma_r(self, r):
        return _sigma_r(self, r)


def EHPower(cosmo, redshift=0.0):
    """No-wiggle Eisenstein-Hu linear power (the reference CLI default)."""
    return LinearPower(cosmo, redshift, "NoWiggleEisensteinHu")
