"""Writes Del2k_planck13.npz, the Planck13 Delta^2(k) table at z = 0 that the
reference computes with cosmolopy and caches as cosmo_files/Del2k.npz
(cosmo_files/sigmas.py:44-52: k = exp(arange(-21, 21, 0.1)) Mpc^-1).
cosmolopy is not available here, so the table is taken from the reference's
own cache (loaded without pickle) and stored as package data.

    python 21cmpy_amd/cosmo_files/make_planck_table.py /root/reference/cosmo_files/Del2k.npz
"""
import os
import sys

import numpy as np


def main(src):
    with np.load(src, allow_pickle=False) as f:
        k, d2 = np.asarray(f["k"], np.float64), np.asarray(f["Del2"], np.float64)
    out = os.path.join(os.path.dirname(os.path.abspath(__file__)), "Del2k_planck13.npz")
    np.savez(out, k=k, Del2=d2)
    print(f"{out}: {k.size} rows, k {k[0]:.3g} .. {k[-1]:.3g} Mpc^-1")


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference/cosmo_files/Del2k.npz")
