"""Planck-cosmology mode of the barrier parameters, without cosmolopy.

The reference obtains the excursion-set barrier from cosmolopy
(find_HII_bubbles.py:57-60):

    sigmamin, deltac = pb.sig_del(ION_Tvir_MIN, Z, **COSMO)
    fc_mean_ps = pb.collapse_fraction(sigmamin, deltac)

and S(R) from its cached Planck13 Delta^2(k) table (cosmo_files/sigmas.py:
44-52, 224-232).  This module restates those quantities from their published
definitions, for COSMO = PLANCK13 (Parameter_files/INIT_PARAMS.py):

* Delta^2(k) at z = 0: the reference's own cached table (Del2k_planck13.npz,
  written by make_planck_table.py), interpolated in log k as the reference's
  interp1d does on its grid; or, analytically, the Eisenstein & Hu (1998,
  ApJ 496, 605, eqs. 29-31) zero-baryon-oscillation fit normalised to
  sigma_8 (``EHSpectrum``; within 3 % of the table, which carries
  baryon oscillations).
* D(z): Carroll, Press & Turner (1992, ARA&A 30, 499) fit for a flat
  Lambda universe, normalised to D(0) = 1 (the approximation cosmolopy's
  fgrowth documents), and the Liddle et al. fit the reference's own C code
  carries (``dicke``, Programs/ps.c:16-45); (dD/dt)/D for the velocities
  at z (Programs/cosmo_functions.py:7-21).
* M_vir(T_vir, z): Barkana & Loeb (2001, Phys. Rep. 349, 125) eq. 26 with
  Delta_c of Bryan & Norman (1998) for a flat universe; mean molecular
  weight 0.59 at T >= 1e4 K (ionized primordial gas), 1.22 below.
* R(M) = (3 M / (4 pi Omega_m rho_crit))^(1/3), rho_crit = 2.775e11 h^2
  M_sun Mpc^-3 (comoving).
* sigma(R) at z = 0 by the reference's trapezoid rule over ln k
  (sigmas.sig0_trapz), and sig_del = (sigma(R(M_vir(T, z))), 1.686 / D(z)).
* collapse_fraction(sigma_min, delta_c) = erfc(delta_c / (sqrt(2) sigma_min)).
"""
import math
import os

import numpy as np

from ..Parameter_files import INIT_PARAMS as P
from .sigmas import sig0_trapz

DELTA_C_LINEAR = 1.686
RHO_CRIT_H2 = 2.775e11  # M_sun Mpc^-3 h^-2 (comoving critical density today)
TABLE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "Del2k_planck13.npz")


def cosmo_params(cosmo=None):
    c = dict(P.COSMO if cosmo is None else cosmo)
    return c


# ------------------------------------------------------------------ spectrum

class TableSpectrum:
    """Delta^2(k) at z = 0 from the cached Planck13 table (k in Mpc^-1),
    interpolated linearly in ln k; zero outside the table."""

    def __init__(self, path=TABLE):
        with np.load(path, allow_pickle=False) as f:
            self.k = np.asarray(f["k"], np.float64)
            self.del2_table = np.asarray(f["Del2"], np.float64)
        self._lk = np.log(self.k)
        self._s0 = {}  # S(R) by radius (the table does not change)

    def del2(self, k):
        k = np.asarray(k, dtype=np.float64)
        return np.interp(np.log(k), self._lk, self.del2_table, left=0.0, right=0.0)

    def sig0(self, R):
        """S(R) by the reference's rule (sigmas.py:176-191, 224-232)."""
        R = float(R)
        s = self._s0.get(R)
        if s is None:
            s = self._s0[R] = sig0_trapz(R, self.del2)
        return s


def eh_nowiggle_transfer(k, cosmo=None, t_cmb=2.728):
    """Eisenstein & Hu (1998) eqs. 29-31: the transfer function without
    baryon oscillations, with their suppression (alpha_Gamma, sound horizon
    fit); k in Mpc^-1."""
    c = cosmo_params(cosmo)
    h, om, ob = c["h"], c["omega_M_0"], c["omega_b_0"]
    theta = t_cmb / 2.7
    om_h2, ob_h2, fb = om * h * h, ob * h * h, ob / om
    s = 44.5 * math.log(9.83 / om_h2) / math.sqrt(1.0 + 10.0 * ob_h2 ** 0.75)
    alpha = 1.0 - 0.328 * math.log(431.0 * om_h2) * fb + 0.38 * math.log(22.3 * om_h2) * fb * fb
    k = np.asarray(k, dtype=np.float64)
    gamma = om * h * (alpha + (1.0 - alpha) / (1.0 + (0.43 * k * s) ** 4))
    q = k * theta * theta / (gamma * h)
    L = np.log(2.0 * math.e + 1.8 * q)
    C = 14.2 + 731.0 / (1.0 + 62.5 * q)
    return L / (L + C * q * q)


class EHSpectrum:
    """Delta^2(k) = A k^(3+n) T(k)^2 with the no-wiggle transfer function,
    A set so that sigma(8/h Mpc) = sigma_8 by the same trapezoid rule."""

    def __init__(self, cosmo=None):
        c = cosmo_params(cosmo)
        self.cosmo = c
        self.n = c["n"]
        self.amp = 1.0
        s8 = sig0_trapz(8.0 / c["h"], self.del2)
        self.amp = c["sigma_8"] ** 2 / s8
        self._s0 = {}  # S(R) by radius, once the amplitude is set

    def del2(self, k):
        k = np.asarray(k, dtype=np.float64)
        return self.amp * k ** (3.0 + self.n) * eh_nowiggle_transfer(k, self.cosmo) ** 2

    def sig0(self, R):
        R = float(R)
        s = self._s0.get(R)
        if s is None:
            s = self._s0[R] = sig0_trapz(R, self.del2)
        return s


_TABLE_SPECTRUM = None


def planck_spectrum(kind="table"):
    global _TABLE_SPECTRUM
    if kind == "table":
        if _TABLE_SPECTRUM is None:
            _TABLE_SPECTRUM = TableSpectrum()
        return _TABLE_SPECTRUM
    if kind == "eh_nowiggle":
        return EHSpectrum()
    raise ValueError(f"unknown spectrum {kind!r}")


def sigma_r(R, spectrum=None):
    """RMS linear density contrast at z = 0 in a top-hat of radius R Mpc."""
    spec = planck_spectrum() if spectrum is None else spectrum
    return math.sqrt(spec.sig0(float(R)))


# ------------------------------------------------------------------ growth

def _omegas(z, cosmo):
    om, ol = cosmo["omega_M_0"], cosmo["omega_lambda_0"]
    a3 = (1.0 + z) ** 3
    e2 = om * a3 + ol + (1.0 - om - ol) * (1.0 + z) ** 2
    return om * a3 / e2, ol / e2


def fgrowth(z, cosmo=None):
    """Linear growth factor D(z), D(0) = 1: Carroll, Press & Turner (1992)
    g(Omega_m, Omega_Lambda) = 2.5 Omega_m / (Omega_m^(4/7) - Omega_Lambda
    + (1 + Omega_m/2)(1 + Omega_Lambda/70)), D = g(z) / ((1+z) g(0))."""
    c = cosmo_params(cosmo)

    def g(zz):
        om, ol = _omegas(zz, c)
        return 2.5 * om / (om ** (4.0 / 7.0) - ol + (1.0 + om / 2.0) * (1.0 + ol / 70.0))

    z = np.asarray(z, dtype=np.float64)
    out = np.vectorize(lambda zz: g(zz) / ((1.0 + zz) * g(0.0)))(z)
    return float(out) if out.ndim == 0 else out


def dicke(z, cosmo=None):
    """The growth fit of the reference's Programs/ps.c:16-45 for a flat
    cosmological-constant universe without radiation (Liddle et al. 1996,
    eqs. 6-8), normalised to 1 at z = 0."""
    c = cosmo_params(cosmo)
    om = c["omega_M_0"]
    ol = c["omega_lambda_0"]

    def d(omz):
        return 2.5 * omz / (1.0 / 70.0 + omz * (209.0 - omz) / 140.0 + omz ** (4.0 / 7.0))

    z = np.asarray(z, dtype=np.float64)
    omz = om * (1.0 + z) ** 3 / (ol + om * (1.0 + z) ** 3)
    out = d(omz) / (d(om) * (1.0 + z))
    return float(out) if out.ndim == 0 else out


H0_PER_H = 1.0e5 / 3.0856775814913673e22  # 100 km/s/Mpc in s^-1


def hubble(z, cosmo=None):
    """H(z) in s^-1."""
    c = cosmo_params(cosmo)
    om, ol = c["omega_M_0"], c["omega_lambda_0"]
    e2 = om * (1.0 + z) ** 3 + ol + (1.0 - om - ol) * (1.0 + z) ** 2
    return c["h"] * H0_PER_H * math.sqrt(e2)


def ddt_growth_over_growth(z, cosmo=None):
    """(dD/dt) / D at z in s^-1 (cosmo_functions.py:7-21 dDdtoverD): dD/dz by a
    central difference, dt/dz = -1 / ((1+z) H(z))."""
    dz = 1.0e-4 * max(1.0, float(z))
    lo = max(float(z) - dz, 0.0)
    ddz = (fgrowth(float(z) + dz, cosmo) - fgrowth(lo, cosmo)) / (float(z) + dz - lo)
    return -ddz * (1.0 + z) * hubble(z, cosmo) / fgrowth(z, cosmo)


# ------------------------------------------------------------------ haloes

def virial_mass(t_vir, z, cosmo=None, mu=None):
    """Halo mass (M_sun) of virial temperature t_vir (K) at redshift z:
    Barkana & Loeb (2001) eq. 26,
    M = 1e8/h (mu/0.6)^(-3/2) (Omega_m/Omega_m^z Delta_c/(18 pi^2))^(-1/2)
        (T/1.98e4 K (1+z)/10)^(3/2),
    Delta_c = 18 pi^2 + 82 d - 39 d^2, d = Omega_m^z - 1 (Bryan & Norman 1998)."""
    c = cosmo_params(cosmo)
    if mu is None:
        mu = 0.59 if t_vir >= 1.0e4 else 1.22
    omz, _ = _omegas(z, c)
    d = omz - 1.0
    delta_c = 18.0 * math.pi ** 2 + 82.0 * d - 39.0 * d * d
    return (1.0e8 / c["h"] * (mu / 0.6) ** -1.5 * (c["omega_M_0"] / omz * delta_c / (18.0 * math.pi ** 2)) ** -0.5
            * (t_vir / 1.98e4 * (1.0 + z) / 10.0) ** 1.5)


def mass_to_radius(mass, cosmo=None):
    """Comoving Lagrangian radius (Mpc) of a mass (M_sun)."""
    c = cosmo_params(cosmo)
    rho = c["omega_M_0"] * RHO_CRIT_H2 * c["h"] ** 2
    return (3.0 * np.asarray(mass, dtype=np.float64) / (4.0 * math.pi * rho)) ** (1.0 / 3.0)


def sig_del(t_vir, z, cosmo=None, spectrum=None):
    """(sigma_min, delta_c) as find_HII_bubbles.py:57 obtains them: the z = 0
    RMS fluctuation on the scale of the smallest ionizing halo and the
    collapse barrier 1.686 / D(z)."""
    r_min = float(mass_to_radius(virial_mass(t_vir, z, cosmo), cosmo))
    return sigma_r(r_min, spectrum), DELTA_C_LINEAR / fgrowth(z, cosmo)


def collapse_fraction(sigma_min, delta_c, sigma_mass=0.0, delta=0.0):
    """erfc((delta_c - delta) / sqrt(2 (sigma_min^2 - sigma_mass^2)))."""
    return math.erfc((delta_c - delta) / math.sqrt(2.0 * (sigma_min ** 2 - sigma_mass ** 2)))


def barrier(z, t_vir=None, cosmo=None, spectrum=None):
    """Everything conv_bubbles needs from cosmology at redshift z:
    {'deltac', 'sigma_min2', 'growth', 'fc_mean', 'sig0': S(R) callable}."""
    t = float(P.ION_Tvir_MIN if t_vir is None else t_vir)
    spec = planck_spectrum() if spectrum is None else spectrum
    smin, dc = sig_del(t, z, cosmo, spec)
    return {"deltac": dc, "sigma_min2": smin * smin, "growth": fgrowth(z, cosmo),
            "fc_mean": collapse_fraction(smin, dc), "sig0": spec.sig0}


def sigma_8(spectrum=None, cosmo=None):
    """sigma(8/h Mpc) of a spectrum (a check of its normalisation)."""
    c = cosmo_params(cosmo)
    return sigma_r(8.0 / c["h"], spectrum)


__all__ = ["TableSpectrum", "EHSpectrum", "eh_nowiggle_transfer", "planck_spectrum", "sigma_r", "fgrowth", "dicke",
           "hubble", "ddt_growth_over_growth", "virial_mass", "mass_to_radius", "sig_del", "collapse_fraction",
           "barrier", "sigma_8"]
