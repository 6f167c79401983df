"""Redshift scroll (drop-in for Programs/Drives.py).

``drive_zscroll_noTs(ZHIGH, ZLOW, DZ)`` follows Drives.py:7-14: the initial
conditions are built once (initialize), then the density field is evolved
to every z of np.arange(ZHIGH, ZLOW, DZ) (perturb_field).  Everything stays
in HBM.  With ``bubbles=True`` the ionization fields of all redshifts are
computed by ONE batched find_HII_bubbles call (the redshifts share HII_DIM
and BOX_LEN, so they form one group with their own barriers 1.686/D(z)),
followed by the brightness temperature and P21(k) of each redshift.

Without ``sigma_min2`` the run follows the Planck13 cosmology of INIT_PARAMS
as the reference does with cosmolopy: the initial conditions drawn from the
Planck13 P(k), D(z) and, per redshift, the barrier and S(R) of the same
spectrum from cosmo_files.planck (ION_Tvir_MIN).  With an explicit ``sigma_min2``
the initial conditions follow the benchmark power law, the growth factors
come from ``growth`` (a callable, default D(z) = 0.12 * 9 / (1 + z), anchored
at D(8) = 0.12) and every redshift uses that sigma_min^2.  The mass move is
cloud-in-cell by default (``scheme``; "ngp" is the reference's truncation).
"""
import numpy as np

from ..Parameter_files import INIT_PARAMS as P
from ..cosmo_files import planck
from . import delta_T, initialize, perturb_field
from .find_HII_bubbles_batched import BatchedBubbles, linear_growth


def drive_zscroll_noTs(ZHIGH=16.0, ZLOW=10.0, DZ=-2.0, DIM=None, HII_DIM=None, BOX_LEN=None, seed=0, growth=None,
                       zeta=None, sigma_min2=None, scheme="cic", bubbles=True, nbins=40, save=False, out_dir=None,
                       quiet=False):
    """Returns {'z', 'density': [HII_DIM^3 tensors]} and, with bubbles,
    'ionized' (tensors), 'x_HI' and 'P21' (per z: (k_edges, k_ave, ps_ave))."""
    DIM = int(P.DIM if DIM is None else DIM)
    HII_DIM = int(P.HII_DIM if HII_DIM is None else HII_DIM)
    L = float(P.BOX_LEN if BOX_LEN is None else BOX_LEN)
    planck_mode = sigma_min2 is None
    grow = growth if growth is not None else (planck.fgrowth if planck_mode else linear_growth)
    zeta = float(P.ZETA if zeta is None else zeta)
    if not quiet:
        print("Now calling initialize")
    # one cosmology throughout: the Planck13 spectrum with the Planck13 barrier,
    # the benchmark power law with an explicit sigma_min^2
    ics = initialize.init(DIM=DIM, HII_DIM=HII_DIM, BOX_LEN=L, seed=seed, pspec="planck" if planck_mode else "powerlaw",
                          save=save, out_dir=out_dir)
    zs = np.arange(ZHIGH, ZLOW, DZ)
    d_init = grow(P.INITIAL_REDSHIFT)
    dens = []
    for z in zs:
        if not quiet:
            print(z)
        dens.append(perturb_field.evolve_zeldovich(z, ics["deltax"], ics["v"], grow(z), d_init, L, scheme,
                                                   save=save, out_dir=out_dir, velocities=False))
    out = {"z": zs, "density": dens}
    if not bubbles:
        return out
    if planck_mode:
        specs = [{"D": HII_DIM, "box_len": L, "zeta": zeta, "z": float(z)} for z in zs]
    else:
        specs = [{"D": HII_DIM, "box_len": L, "zeta": zeta, "deltac": 1.686 / grow(z), "sigma_min2": sigma_min2}
                 for z in zs]
    bb = BatchedBubbles(specs, fil=P.HII_FILTER, LE=False, graph=False)
    bb.set_deltas(dens)
    bb.run()
    out["ionized"] = [bb.ionized(i).clone() for i in range(len(zs))]
    out["x_HI"] = bb.x_HI()
    edges = delta_T.log_edges(L, HII_DIM, nbins)
    out["P21"] = []
    for z, ion, d in zip(zs, out["ionized"], dens):
        dT, s = delta_T.brightness_temperature((1.0 - ion).contiguous(), d, float(z))
        K, k_ave, ps_ave, _ = delta_T.power_spectrum(dT, L, edges, field_sum=s)
        out["P21"].append((K, k_ave, ps_ave))
    return out


if __name__ == "__main__":
    drive_zscroll_noTs(16.0, 10.0, -2.0)
