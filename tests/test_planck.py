"""Planck-cosmology mode of the barrier (cosmo_files/planck.py) against the
reference's own tables and identities (CPU).

Pinned:
* the Planck13 Delta^2(k) table: its sigma_8 by the S(R) rule equals COSMO's
  0.829 (cosmo_files/Del2k.npz, Parameter_files/INIT_PARAMS.py);
* S(R) of the table against the reference's sig0.npz (tests/golden/
  planck_tables.npz, every 10th row), and the mass-radius relation against
  the same rows;
* D(z) against the growth fit of the reference's Programs/ps.c:16-45 (for a
  flat universe the Carroll-Press-Turner form and the Liddle form are the
  same function), D(0) = 1 and the matter-dominated limit D ~ 1/(1+z);
* the virial mass at the point where Barkana & Loeb's formula reduces to its
  prefactor.
Parity unpinned: the exact cosmolopy sig_del (cosmolopy is not installed);
the no-wiggle Eisenstein & Hu spectrum is checked only to 3 % of the table.
"""
import math
import os

import numpy as np
import pytest

from conftest import pkg

GOLDEN = os.path.join(os.path.dirname(__file__), "golden")


def test_table_sigma8_matches_cosmology():
    pl = pkg("cosmo_files.planck")
    assert pl.sigma_8() == pytest.approx(0.829, rel=1e-3)


def test_table_sig0_and_mass_radius_match_reference():
    pl = pkg("cosmo_files.planck")
    t = np.load(os.path.join(GOLDEN, "planck_tables.npz"))
    R = pl.mass_to_radius(t["mass"])
    assert np.max(np.abs(R / t["radius"] - 1)) < 1e-4
    spec = pl.planck_spectrum()
    S = np.array([spec.sig0(r) for r in t["radius"]])
    rel = np.abs(S / t["sig0"] - 1)
    # the reference interpolates the table linearly in k, this module in ln k
    assert rel.max() < 3e-3 and np.median(rel) < 1.5e-3


def test_growth_matches_reference_fit_and_limits():
    pl = pkg("cosmo_files.planck")
    z = np.array([0.0, 0.5, 3.0, 6.0, 8.0, 10.0, 12.0, 20.0, 50.0])
    assert np.allclose(pl.fgrowth(z), pl.dicke(z), rtol=1e-12, atol=0)
    assert pl.fgrowth(0.0) == pytest.approx(1.0, abs=1e-15)
    # matter domination: (1+z) D(z) tends to a constant
    a, b = (1 + 500.0) * pl.fgrowth(500.0), (1 + 1000.0) * pl.fgrowth(1000.0)
    assert a == pytest.approx(b, rel=1e-5)
    # Lambda suppresses growth today: D(z) (1+z) > 1 at z > 0, increasing with z
    g = (1 + z[1:]) * pl.fgrowth(z[1:])
    assert np.all(g > 1) and np.all(np.diff(g) > 0)


def test_virial_mass_prefactor_and_scaling():
    pl = pkg("cosmo_files.planck")
    c = pl.cosmo_params()
    # Omega_m^z -> 1 at high z, so Delta_c -> 18 pi^2: M = 1e8/h (Omega_m)^(-1/2) at T = 1.98e4 K, (1+z) = 10
    # shifted to z where Omega_m^z ~ 1 - 1e-9 via the (1+z)^(3/2) factor
    z = 999.0
    m = pl.virial_mass(1.98e4, z, mu=0.6)
    expect = 1.0e8 / c["h"] * c["omega_M_0"] ** -0.5 * (1000.0 / 10.0) ** 1.5
    assert m == pytest.approx(expect, rel=1e-6)
    # M ~ T^(3/2) at fixed z and mu
    assert pl.virial_mass(4e4, 10.0, mu=0.59) / pl.virial_mass(2e4, 10.0, mu=0.59) == pytest.approx(2 ** 1.5)
    # ionized gas above 1e4 K: mu 0.59, neutral below: 1.22
    assert pl.virial_mass(1.0e4, 10.0) == pytest.approx(pl.virial_mass(1.0e4, 10.0, mu=0.59))


def test_eh_nowiggle_within_three_percent_of_table():
    pl = pkg("cosmo_files.planck")
    tab, eh = pl.planck_spectrum("table"), pl.planck_spectrum("eh_nowiggle")
    assert pl.sigma_8(eh) == pytest.approx(0.829, rel=1e-9)
    k = np.exp(np.linspace(np.log(1e-3), np.log(10.0), 60))
    assert np.max(np.abs(eh.del2(k) / tab.del2(k) - 1)) < 0.03


def test_barrier_and_collapse_fraction():
    pl = pkg("cosmo_files.planck")
    b = pl.barrier(10.0)
    smin = math.sqrt(b["sigma_min2"])
    assert b["deltac"] == pytest.approx(1.686 / pl.fgrowth(10.0))
    assert b["fc_mean"] == pytest.approx(math.erfc(b["deltac"] / (math.sqrt(2) * smin)))
    # the smallest ionizing halo at 1e4 K and z = 10: ~1e8 M_sun, sigma ~ 6 at z = 0
    m = pl.virial_mass(1.0e4, 10.0)
    assert 5e7 < m < 2e8 and 5.0 < smin < 7.0
    # the collapsed fraction grows towards lower z
    assert pl.barrier(8.0)["fc_mean"] > b["fc_mean"] > pl.barrier(12.0)["fc_mean"]


@pytest.mark.parametrize("D,L,z", [(256, 300.0, 8.0), (256, 300.0, 12.0), (128, 200.0, 10.0), (64, 100.0, 9.0)])
def test_planck_mode_host_setup(D, L, z):
    """The Planck-mode scalars give a valid radius loop (every barrier
    denominator real) for the benchmark box sizes."""
    fb = pkg("Programs.find_HII_bubbles")
    pl = pkg("cosmo_files.planck")
    b = pl.barrier(z)
    radii, denoms, fc_mean = fb.bubble_setup(D, L, None, b["deltac"], b["sigma_min2"], b["sig0"])
    assert len(radii) > 10 and np.all(np.isfinite(denoms)) and np.all(denoms > 0)
    assert fc_mean == pytest.approx(np.float32(b["fc_mean"]), rel=1e-6)


def test_growth_rate_limits():
    """(dD/dt)/D = f(z) H(z): f -> 1 in matter domination; today f close to
    Omega_m^0.55 (the CPT fit's derivative, as the reference's finite
    difference of fgrowth gives it, cosmo_functions.py:7-21)."""
    pl = pkg("cosmo_files.planck")
    c = pl.cosmo_params()
    assert pl.ddt_growth_over_growth(1000.0) / pl.hubble(1000.0) == pytest.approx(1.0, abs=1e-5)
    assert pl.ddt_growth_over_growth(8.0) / pl.hubble(8.0) == pytest.approx(1.0, abs=3e-3)
    assert pl.ddt_growth_over_growth(0.0) / pl.hubble(0.0) == pytest.approx(c["omega_M_0"] ** 0.55, rel=0.05)
    # H0 = 100 h km/s/Mpc
    assert pl.hubble(0.0) == pytest.approx(c["h"] * 100.0 / 3.0856775814913673e19, rel=1e-12)
