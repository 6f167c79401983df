"""Writes the committed fixtures of tests/golden.

* reference_tables.npz: a subset of the reference-held tables
  cosmo_files/Del2k.npz (the Delta^2(k) rows whose k the S(R) rule touches for
  R in [0.5, 30] Mpc) and cosmo_files/sig0.npz (every 20th radius in that
  range with its S(R)).  Provenance: yunfanz/21cmpy, cosmo_files/, read with
  numpy.load (no pickles).
* bubbles_D16.npz / bubbles_D32_nl.npz: seeded GRF boxes with the oracle's
  float32 ionized field (regression vectors for the oracle itself).
* planck_tables.npz: every 10th row of cosmo_files/sig0.npz (mass in M_sun,
  radius in Mpc, S(R)) for the Planck-cosmology mode (mass-radius relation and
  S(R) of the Planck13 table).  Provenance as above.

Run from the repository root: python tests/golden/make_golden.py [planck]
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
REF = "/root/reference/cosmo_files"

from oracle import sigma, grf, bubbles  # noqa: E402


def tables():
    d = np.load(os.path.join(REF, "Del2k.npz"))
    s = np.load(os.path.join(REF, "sig0.npz"))
    k, d2 = d["k"], d["Del2"]
    f = sigma.table_del2(k, d2)
    lo, hi = np.inf, -np.inf
    for R in (0.5, 30.0):
        logk = np.arange(-20.0, 20.0, 0.1)
        a, b = sigma._klims(f(np.exp(logk)) * sigma.tophat_window(R * np.exp(logk)) ** 2, logk)
        lo, hi = min(lo, a), max(hi, b)
    keep = (k >= np.exp(lo) * 0.8) & (k <= np.exp(hi) * 1.25)
    i0, i1 = np.nonzero(keep)[0][[0, -1]]
    sel = slice(max(i0 - 1, 0), min(i1 + 2, k.size))
    R = s["radius"]
    rsel = np.nonzero((R >= 0.5) & (R <= 30.0))[0][::20]
    np.savez(os.path.join(HERE, "reference_tables.npz"), k=k[sel], Del2=d2[sel],
             radius=R[rsel], sig0=s["sig0"][rsel])


def bubble_vectors():
    for name, D, L, LE, smin in (("bubbles_D16.npz", 16, 25.0, True, 0.5),
                                 ("bubbles_D32_nl.npz", 32, 50.0, False, 2.5)):
        spec = grf.grf_half_spectrum(D, L, seed=7)
        delta = grf.c2r(spec, D).astype(np.float32)
        radii, _ = sigma.radius_schedule(L, D)
        den = sigma.denominators(radii, smin, lambda R: sigma.sig0_discrete(R, D, L))
        growth = 0.5
        deltac = 1.686 if LE else 1.686 / growth
        if not LE:
            delta = (delta * np.float32(growth)).astype(np.float32)
        fc_mean = np.float32(__import__("scipy.special").special.erfc(deltac / np.sqrt(2 * smin)))
        ion = bubbles.find_bubbles_f32(delta, L, radii, den, 30.0, deltac, LE=LE,
                                       growth_inv=np.float32(deltac / 1.686), fc_mean=fc_mean)
        np.savez_compressed(os.path.join(HERE, name), delta=delta, radii=radii, denoms=den, ionized=ion,
                            box_len=L, sigma_min2=smin, deltac=deltac, zeta=30.0, LE=LE, fc_mean=fc_mean)


def planck_tables():
    s = np.load(os.path.join(REF, "sig0.npz"))
    np.savez(os.path.join(HERE, "planck_tables.npz"), mass=s["mass"][::10], radius=s["radius"][::10],
             sig0=s["sig0"][::10])


if __name__ == "__main__":
    if sys.argv[1:] == ["planck"]:
        planck_tables()
    else:
        tables()
        bubble_vectors()
        planck_tables()
    print("ok")
