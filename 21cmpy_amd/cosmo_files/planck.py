"""Planck-cosmology mode of the barrier parameters, without cosmolopy.

The reference obtains the excursion-set barrier from cosmolopy
(find_HII_bubbles.py:57-60):

    sigmamin, deltac = pb.sig_del(ION_Tvir_MIN, Z, **COSMO)
    fc_mean_ps = pb.collapse_fraction(sigmamin, deltac)

its S(R) from a cached Planck13 Delta^2(k) table (cosmo_files/sigmas.py:
44-52, 224-232) and its initial-condition spectrum from
pb.power_spectrum(K, 0.0, **COSMO) (initialize.py:27-31).  This module
restates those quantities from their published definitions, for COSMO =
PLANCK13 (Parameter_files/INIT_PARAMS.py):

* P(k) at z = 0: Eisenstein & Hu (1999, ApJ 511, 5) transfer function for
  cold dark matter, baryons and massive neutrinos (``eh99_transfer``), the fit
  cosmolopy wraps, evaluated as the authors' public fitting code does it
  (T_cmb = 2.728 K, an exponent 1.11 in the C(q) term, the baryon and
  neutrino densities floored at Omega = 1e-5 and at least one neutrino
  species, so that with N_nu = 0 a vanishing neutrino fraction 1e-5/Omega_m
  enters), times k^n and normalised to sigma_8 by an adaptive quadrature of
  the top-hat variance at 8/h Mpc (``EH99Spectrum``).  Against the
  reference's cached Del2k.npz the shape agrees to 6e-7 and the amplitude to
  8e-5.  S(R) follows the reference's rule on top of it: Delta^2 tabulated on
  k = exp(arange(-21, 21, 0.1)), interpolated linearly in k, trapezoid rule
  over ln k (sigmas.sig0_trapz).  The Eisenstein & Hu (1998, ApJ 496, 605,
  eqs. 29-31) zero-baryon-oscillation fit is kept as ``EHSpectrum``.
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

import numpy as np
from scipy import integrate

from ..Parameter_files import INIT_PARAMS as P
from .sigmas import sig0_trapz, tophat_window

DELTA_C_LINEAR = 1.686
RHO_CRIT_H2 = 2.775e11  # M_sun Mpc^-3 h^-2 (comoving critical density today)
T_CMB = 2.728  # K, as in the Eisenstein & Hu fitting code
# k grid of the reference's cached Delta^2(k) (cosmo_files/sigmas.py:50)
REFERENCE_LOGK = np.arange(-21.0, 21.0, 0.1)


def cosmo_params(cosmo=None):
    c = dict(P.COSMO if cosmo is None else cosmo)
    return c


# ------------------------------------------------------------------ spectrum

def eh99_transfer(k, cosmo=None, t_cmb=T_CMB):
    """Eisenstein & Hu (1999) transfer function of the cold dark matter plus
    baryon component at z = 0 (k in Mpc^-1).

    Equation numbers refer to the paper: z_eq (1), z_d (2), y_d (3), the sound
    horizon fit s (4), q (5), p_c and p_cb (11), the growth factor with
    free-streaming (10, 12-14), the small-scale suppression alpha_nu (15), the
    effective shape Gamma_eff (16), the suppressed transfer function (17-21)
    and the neutrino master correction (22-24).
    """
    c = cosmo_params(cosmo)
    h = float(c["h"])
    om = float(c["omega_M_0"])
    ol = float(c.get("omega_lambda_0", 1.0 - om))
    ob = max(float(c["omega_b_0"]), 1.0e-5)
    onu = max(float(c.get("omega_n_0", 0.0)), 1.0e-5)
    n_nu = max(float(c.get("N_nu", 0)), 1.0)
    theta2 = (t_cmb / 2.7) ** 2
    w_m, w_b = om * h * h, ob * h * h
    f_b, f_nu = ob / om, onu / om
    f_c = 1.0 - f_b - f_nu
    f_cb, f_nub = f_c + f_b, f_nu + f_b

    one_z_eq = 2.5e4 * w_m / (theta2 * theta2)
    d1 = 0.313 * w_m ** -0.419 * (1.0 + 0.607 * w_m ** 0.674)
    d2 = 0.238 * w_m ** 0.223
    z_d = 1291.0 * w_m ** 0.251 / (1.0 + 0.659 * w_m ** 0.828) * (1.0 + d1 * w_b ** d2)
    y_d = one_z_eq / (1.0 + z_d)
    s = 44.5 * math.log(9.83 / w_m) / math.sqrt(1.0 + 10.0 * w_b ** 0.75)
    p_c = 0.25 * (5.0 - math.sqrt(1.0 + 24.0 * f_c))
    p_cb = 0.25 * (5.0 - math.sqrt(1.0 + 24.0 * f_cb))

    ok = 1.0 - om - ol  # curvature
    e2 = ol + (ok + om)  # at z = 0
    om_z, ol_z = om / e2, ol / e2
    g0 = one_z_eq * 2.5 * om_z / (om_z ** (4.0 / 7.0) - ol_z + (1.0 + om_z / 2.0) * (1.0 + ol_z / 70.0))

    a_nu = (f_c / f_cb * (5.0 - 2.0 * (p_c + p_cb)) / (5.0 - 4.0 * p_cb) * (1.0 + y_d) ** (p_cb - p_c)
            * (1.0 + f_nub * (-0.553 + 0.126 * f_nub * f_nub))
            / (1.0 - 0.193 * math.sqrt(f_nu * n_nu) + 0.169 * f_nu * n_nu ** 0.2)
            * (1.0 + 0.5 * (p_c - p_cb) * (1.0 + 1.0 / ((3.0 - 4.0 * p_c) * (7.0 - 4.0 * p_cb))) / (1.0 + y_d)))
    sqrt_a = math.sqrt(a_nu)
    beta_c = 1.0 / (1.0 - 0.949 * f_nub)

    k = np.asarray(k, dtype=np.float64)
    q = k * theta2 / w_m
    y_fs = 17.2 * f_nu * (1.0 + 0.488 * f_nu ** (-7.0 / 6.0)) * (n_nu * q / f_nu) ** 2
    growth_cb = (1.0 + (g0 / (1.0 + y_fs)) ** 0.7) ** (p_cb / 0.7) * g0 ** (1.0 - p_cb)
    gamma_eff = w_m * (sqrt_a + (1.0 - sqrt_a) / (1.0 + (0.43 * k * s) ** 4))
    q_eff = k * theta2 / gamma_eff
    L = np.log(2.71828 + 1.84 * beta_c * sqrt_a * q_eff)
    C = 14.4 + 325.0 / (1.0 + 60.5 * q_eff ** 1.11)
    t_sup = L / (L + C * q_eff * q_eff)
    q_nu = 3.92 * q * math.sqrt(n_nu / f_nu)
    master = 1.0 + 1.2 * f_nu ** 0.64 * n_nu ** (0.3 + 0.6 * f_nu) / (q_nu ** -1.6 + q_nu ** 0.8)
    return t_sup * master * growth_cb / g0


def eh_nowiggle_transfer(k, cosmo=None, t_cmb=T_CMB):
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


def sigma8_amplitude(shape_fn, cosmo=None):
    """A such that A * shape(k) (a Delta^2(k) shape) has sigma(8/h Mpc) =
    sigma_8, by adaptive quadrature over ln k."""
    c = cosmo_params(cosmo)
    R8 = 8.0 / c["h"]

    def f(lk):
        kk = math.exp(lk)
        return float(shape_fn(np.array([kk]))[0] * tophat_window(np.array([kk * R8]))[0] ** 2)

    I, _ = integrate.quad(f, math.log(1e-7), math.log(1e3), limit=4000, epsabs=0.0, epsrel=1e-11)
    return c["sigma_8"] ** 2 / I


class _Spectrum:
    """Delta^2(k) = amp * shape(k) with the reference's S(R) rule on top."""

    def _init(self, cosmo):
        self.cosmo = cosmo_params(cosmo)
        self.n = float(self.cosmo["n"])
        self.amp = sigma8_amplitude(self.del2_shape, self.cosmo)
        self._ktab = np.exp(REFERENCE_LOGK)
        self._dtab = self.del2_exact(self._ktab)
        self._s0 = {}  # S(R) by radius

    def del2_exact(self, k):
        return self.amp * self.del2_shape(k)

    def del2(self, k):
        """Delta^2 tabulated on the reference's k grid and interpolated
        linearly in k, as interp1d at cosmo_files/sigmas.py:50-52."""
        return np.interp(np.asarray(k, dtype=np.float64), self._ktab, self._dtab)

    def pk(self, k):
        """P(k) in Mpc^3 at z = 0 (k in Mpc^-1), 2 pi^2 Delta^2 / k^3."""
        k = np.asarray(k, dtype=np.float64)
        with np.errstate(divide="ignore", invalid="ignore"):
            return np.where(k > 0, 2.0 * math.pi ** 2 * self.del2_exact(k) / np.where(k > 0, k, 1.0) ** 3, 0.0)

    def init_table(self, DIM, box_len, n=512):
        """(K, P) float32 tables of initialize.py:27-31: n points log-spaced
        from DELTA_K/10 to DELTA_K sqrt(3) DIM, DELTA_K = float32(2 pi/BOX_LEN)."""
        dk = float(np.float32(2.0 * np.pi / float(np.float32(box_len))))
        K = np.logspace(np.log10(dk / 10.0), np.log10(dk * np.sqrt(3.0) * int(DIM)), int(n))
        return K.astype(np.float32), self.pk(K).astype(np.float32)

    def sig0(self, R):
        """S(R) by the reference's rule (sigmas.py:176-191, 224-232)."""
        R = float(R)
        s = self._s0.get(R)
        if s is None:
            s = self._s0[R] = sig0_trapz(R, self.del2)
        return s


class EH99Spectrum(_Spectrum):
    """Planck13 Delta^2(k) = A k^(3+n) T(k)^2 with the Eisenstein & Hu (1999)
    transfer function, sigma_8-normalised (the spectrum of initialize.py:30
    and of the reference's cached S(R) table)."""

    def __init__(self, cosmo=None):
        self.cosmo = cosmo_params(cosmo)
        self._init(self.cosmo)

    def del2_shape(self, k):
        k = np.asarray(k, dtype=np.float64)
        return k ** (3.0 + float(self.cosmo["n"])) * eh99_transfer(k, self.cosmo) ** 2


class EHSpectrum(_Spectrum):
    """Delta^2(k) = A k^(3+n) T(k)^2 with the Eisenstein & Hu (1998)
    no-wiggle transfer function, sigma_8-normalised."""

    def __init__(self, cosmo=None):
        self.cosmo = cosmo_params(cosmo)
        self._init(self.cosmo)

    def del2_shape(self, k):
        k = np.asarray(k, dtype=np.float64)
        return k ** (3.0 + float(self.cosmo["n"])) * eh_nowiggle_transfer(k, self.cosmo) ** 2


_SPECTRA = {}


def planck_spectrum(kind="eh99"):
    """The Planck13 spectrum object (cached): "eh99" (default) or "eh_nowiggle"."""
    if kind not in ("eh99", "eh_nowiggle"):
        raise ValueError(f"unknown spectrum {kind!r}")
    if kind not in _SPECTRA:
        _SPECTRA[kind] = EH99Spectrum() if kind == "eh99" else EHSpectrum()
    return _SPECTRA[kind]


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


__all__ = ["EH99Spectrum", "EHSpectrum", "eh99_transfer", "eh_nowiggle_transfer", "sigma8_amplitude",
           "planck_spectrum", "sigma_r", "fgrowth", "dicke",
           "hubble", "ddt_growth_over_growth", "virial_mass", "mass_to_radius", "sig_del", "collapse_fraction",
           "barrier", "sigma_8"]
