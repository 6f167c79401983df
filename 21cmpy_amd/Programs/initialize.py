"""Initial conditions (drop-in for Programs/initialize.py).

``init`` follows initialize.init (initialize.py:33-111): a Gaussian random
field on the DIM^3 grid, the real-space box, the box smoothed with the
type-0 window at R = L_FACTOR*BOX_LEN/HII_DIM and nearest-picked to HII_DIM,
and the three smoothed, subsampled Zel'dovich displacement boxes.  The field
is drawn directly as a Hermitian half spectrum on the GPU (cm21_grf_init);
everything stays in device memory and the .npy files are written only when
``save`` is set, under the reference's names (initialize.py:82,93,109).

The power spectrum is, by default, the reference's: the Planck13
(INIT_PARAMS.COSMO) Eisenstein & Hu P(k) at z = 0 tabulated at 512 points
log-spaced from DELTA_K/10 to DELTA_K sqrt(3) DIM (initialize.py:27-31) and
interpolated linearly per mode on the GPU (cm21_grf_init_table), with the
spectrum computed by cosmo_files.planck instead of cosmolopy.
``pspec="powerlaw"`` gives the benchmark configurations' P(k) = A k^n
exp(-(k/k_nyq)^2) with A set by the per-cell variance
(cosmo_files.sigmas.PowerLawSpectrum).
"""
import os

import numpy as np
import torch

from .. import gpu, _lib
from ..Parameter_files import INIT_PARAMS as P
from ..cosmo_files import planck
from ..cosmo_files.sigmas import power_law_spectrum

BLOCK_VOLUME = 512  # points of the P(k) table (initialize.py:28)


def smoothing_radius(box_len, hii_dim):
    """smoothR = float32(L_FACTOR*BOX_LEN/HII_DIM) (initialize.py:87,99)."""
    return np.float32(float(np.float32(P.L_FACTOR) * np.float32(box_len)) / float(hii_dim))


def box_names(DIM, HII_DIM, BOX_LEN):
    L = float(BOX_LEN)
    return {
        "deltax": "deltax_z0.00_{0:d}_{1:.0f}Mpc".format(int(DIM), L),
        "smoothed": "smoothed_deltax_z0.00_{0:d}_{1:.0f}Mpc".format(int(HII_DIM), L),
        "v": ["v{0}overddot_{1:d}_{2:.0f}Mpc".format(m, int(HII_DIM), L) for m in "xyz"],
    }


def pk_table(DIM, BOX_LEN, spectrum=None):
    """float32 (K, P) table of init_pspec (initialize.py:27-31) for a spectrum
    object (default: the Planck13 spectrum of INIT_PARAMS.COSMO)."""
    sp = planck.planck_spectrum() if spectrum is None else spectrum
    return sp.init_table(int(DIM), float(BOX_LEN), BLOCK_VOLUME)


def init(DIM=None, HII_DIM=None, BOX_LEN=None, seed=0, box_id=0, pspec="planck", sigma_cell2=0.36, n_index=-2.5,
         with_velocity=True, save=True, out_dir=None):
    """Returns {'spec', 'deltax', 'smoothed', 'v': (vx, vy, vz)} as CUDA tensors
    (and 'pk_table' for a tabulated spectrum).

    pspec: "planck" (the reference's spectrum, default), "powerlaw" (the
    benchmark recipe with sigma_cell2 and n_index), or an object with
    ``init_table(DIM, BOX_LEN, n)`` returning float32 (K, P)."""
    _lib.require_gpu()
    DIM = int(P.DIM if DIM is None else DIM)
    HII_DIM = int(P.HII_DIM if HII_DIM is None else HII_DIM)
    L = float(P.BOX_LEN if BOX_LEN is None else BOX_LEN)
    if DIM % HII_DIM:
        raise ValueError(f"HII_DIM={HII_DIM} must divide DIM={DIM}")
    table = None
    if isinstance(pspec, str) and pspec == "powerlaw":
        spectrum = power_law_spectrum(DIM, L, sigma_cell2, n_index)
        spec = gpu.grf_init(DIM, L, spectrum.amp, seed, box_id, n_index, spectrum.k_cut)
    else:
        if isinstance(pspec, str) and pspec != "planck":
            raise ValueError(f"pspec must be 'planck', 'powerlaw' or a spectrum object, not {pspec!r}")
        table = pk_table(DIM, L, None if isinstance(pspec, str) else pspec)
        spec = gpu.grf_init_table(DIM, L, table[0], table[1], seed, box_id)
    deltax = gpu.c2r(spec)                                   # initialize.py:79-82
    R = smoothing_radius(L, HII_DIM)
    work = gpu.hii_filter(spec, L, R, filter_type=0, ref_fallthrough=True, korder=1)
    large = gpu.c2r(work, preserve_input=False)              # initialize.py:86-93
    smoothed = gpu.subsample(large, HII_DIM)
    out = {"spec": spec, "deltax": deltax, "smoothed": smoothed}
    if table is not None:
        out["pk_table"] = table
    if with_velocity:
        vs = []
        for comp in range(3):                                # initialize.py:102-109
            gpu.velocity(spec, L, comp, filter_type=0, R=R, out=work)
            gpu.c2r(work, out=large, preserve_input=False)
            vs.append(gpu.subsample(large, HII_DIM))
        out["v"] = tuple(vs)
    del work, large
    if save:
        d = os.path.join(out_dir or P.IO_DIR, "Boxes")
        os.makedirs(d, exist_ok=True)
        names = box_names(DIM, HII_DIM, L)
        np.save(os.path.join(d, names["deltax"]), deltax.cpu().numpy())
        np.save(os.path.join(d, names["smoothed"]), smoothed.cpu().numpy())
        if with_velocity:
            for n, v in zip(names["v"], out["v"]):
                np.save(os.path.join(d, n), v.cpu().numpy())
    return out


def run(**kwargs):
    """initialize.py:283-298 without the out-of-core branch: HBM holds the
    whole DIM^3 problem."""
    torch.cuda.synchronize()
    return init(**kwargs)
