"""21-cm brightness temperature and power spectrum (drop-in for Programs/delta_T.py).

``brightness_temperature`` is delta_T.py:73-76 on the GPU
(dT = c(z) xH (1 + delta) in mK, FP64 box sum for the mean).
``power_spectrum`` is delta_T.py:80-98 with the binning of delta_T.cu:5-29:
Delta^2(k) = k^3 |d_k|^2 / (2 pi^2 V) of d = (dT/<dT> - 1) V/N, averaged per
bin with the bin of a mode being the first edge K[c] >= |k|.  Corrections to
the reference (which passes DIM for HII_DIM, reads complex data as float and
races on its sums): the true box size, DC excluded, modes above the last edge
dropped, FP64 sums and exact counts.
"""
import os

import numpy as np
import torch

from .. import gpu, _lib
from ..Parameter_files import INIT_PARAMS as P


def tb_constant(z, cosmo=None):
    """27 (Ob h^2/0.023) sqrt((0.15/Om/h^2)(1+z)/10) in float32 (delta_T.py:73-74)."""
    c = P.COSMO if cosmo is None else cosmo
    return np.float32(27 * (c["omega_b_0"] * c["h"] * c["h"] / 0.023) *
                      np.sqrt((0.15 / c["omega_M_0"] / c["h"] / c["h"]) * (1 + z) / 10.0))


def log_edges(box_len, hii_dim, nbins):
    """nbins log-spaced upper bin edges from k_fund = 2 pi/L to k_nyq = pi D/L."""
    return np.logspace(np.log10(2 * np.pi / box_len), np.log10(np.pi * hii_dim / box_len), int(nbins))


def brightness_temperature(xH, delta, z, cosmo=None):
    """Returns (dT, FP64 device sum of dT)."""
    _lib.require_gpu()
    return gpu.delta_T(xH, delta, tb_constant(z, cosmo))


def power_spectrum(field, box_len, edges, field_sum=None, dimensional=None):
    """(K, k_ave, ps_ave, counts) as NumPy arrays, delta_T.py:87-103."""
    dim = P.DIMENSIONAL_T_POWER_SPEC if dimensional is None else dimensional
    ps, ks, cnt = gpu.power_spectrum(field, box_len, edges, field_sum=field_sum, dimensional=dim)
    ps, ks, cnt = ps.cpu().numpy(), ks.cpu().numpy(), cnt.cpu().numpy()
    with np.errstate(invalid="ignore", divide="ignore"):
        k_ave = np.where(cnt > 0, ks / np.maximum(cnt, 1), 0.0)
        ps_ave = np.where(cnt > 0, ps / np.maximum(cnt, 1), 0.0)
    return np.asarray(edges), k_ave, ps_ave, cnt


def delta_T_name(Z, nf, eff, ave, hii_dim, box_len):
    """delta_T.py:77-78."""
    return ("delta_T_no_halos_z{0:.2f}_nf{1:f}_useTs{2:d}_zetaX{3:.1e}_TvirminX{4:.1e}_aveTb{5:.2f}_{6:d}_{7:d}Mpc.npy"
            .format(Z, nf, P.USE_TS_IN_21CM, eff, P.ION_Tvir_MIN, ave, int(hii_dim), int(box_len)))


def run(xHfile, IO_DIR=None, nbins=40, delta_file=None):
    """delta_T.py:32-103: reads an xH box and the matching updated density,
    writes the dT box and returns (K, k_ave, ps_ave)."""
    from ..tocmfastpy import boxio
    io = IO_DIR if IO_DIR is not None else P.IO_DIR
    xH = np.load(xHfile)
    p = boxio.parse_filename(os.path.splitext(xHfile)[0])
    Z, hii, L = p["z"], p["dim"], float(p["BoxSize"])
    if delta_file is None:
        # the reference reads "..._z0{:.2f}..." (delta_T.py:56) while perturb_field
        # writes "..._z{:.2f}..." (perturb_field.py:108); accept both
        cands = ["updated_smoothed_deltax_z{0:.2f}_{1:d}_{2:.0f}Mpc.npy".format(Z, hii, L),
                 "updated_smoothed_deltax_z0{0:.2f}_{1:d}_{2:.0f}Mpc.npy".format(Z, hii, L)]
        delta_file = next((os.path.join(io, "Boxes", c) for c in cands if os.path.exists(os.path.join(io, "Boxes", c))),
                          None)
        if delta_file is None:
            raise FileNotFoundError(f"no updated density box for z={Z:.2f}, dim={hii}, L={L:.0f} under {io}/Boxes")
    delta = np.load(delta_file)
    xt = torch.from_numpy(np.ascontiguousarray(xH, np.float32)).cuda()
    dt = torch.from_numpy(np.ascontiguousarray(delta, np.float32)).cuda()
    dT, s = brightness_temperature(xt, dt, Z)
    ave = float(s.item()) / dT.numel()
    os.makedirs(os.path.join(io, "Boxes"), exist_ok=True)
    np.save(os.path.join(io, "Boxes", delta_T_name(Z, p.get("nf", 0.0), p.get("eff", P.ZETA), ave, hii, L)),
            dT.cpu().numpy())
    K, k_ave, ps_ave, _ = power_spectrum(dT, L, log_edges(L, hii, nbins), field_sum=s)
    return K, k_ave, ps_ave
