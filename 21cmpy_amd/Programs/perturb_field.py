"""Density field at redshift z (drop-in for Programs/perturb_field.py).

``evolve_zeldovich`` follows perturb_field.py:65-108: the displacement boxes
are scaled by (D(z) - D_init)/BOX_LEN, the DIM^3 particles move and deposit
(1 + D_init delta) on the HII grid, and the result is divided by MASS_FACTOR
minus one.  ``scheme='ngp'`` is the reference's truncation; ``'cic'``
(default) deposits cloud-in-cell.  The velocity recompute of :111-126 feeds
only redshift-space distortions and is not part of this stage.
``evolve_linear`` is perturb_field.py:39-50.  Growth factors are arguments
(cosmolopy's fgrowth is not available).
"""
import os

import numpy as np
import torch

from .. import gpu, _lib
from ..Parameter_files import INIT_PARAMS as P
from ..cosmo_files import planck
from . import initialize


def displacement_scale(growth, growth_init, box_len):
    """(fgrowth - primordial_fgrowth)/BOX_LEN in float32 (perturb_field.py:93-95)."""
    return np.float32(np.float32(np.float32(growth) - np.float32(growth_init)) / np.float32(box_len))


def updated_name(z, HII_DIM, BOX_LEN):
    return "updated_smoothed_deltax_z{0:.2f}_{1:d}_{2:.0f}Mpc".format(float(z), int(HII_DIM), float(BOX_LEN))


def evolve_zeldovich(z, deltax, v, growth, growth_init, box_len=None, scheme="cic", save=False, out_dir=None):
    """deltax: DIM^3 CUDA tensor; v: (vx, vy, vz) HII_DIM^3 CUDA tensors."""
    _lib.require_gpu()
    L = float(P.BOX_LEN if box_len is None else box_len)
    s = displacement_scale(growth, growth_init, L)
    upd = gpu.move_mass(deltax, v[0], v[1], v[2], s, np.float32(growth_init), scheme)
    if save:
        d = os.path.join(out_dir or P.IO_DIR, "Boxes")
        os.makedirs(d, exist_ok=True)
        np.save(os.path.join(d, updated_name(z, upd.shape[0], L)), upd.cpu().numpy())
    return upd


def evolve_linear(z, deltax, growth, box_len=None, save=False, out_dir=None):
    """delta * D(z) (perturb_field.py:39-50)."""
    L = float(P.BOX_LEN if box_len is None else box_len)
    t = deltax if isinstance(deltax, torch.Tensor) else torch.from_numpy(np.asarray(deltax, np.float32)).cuda()
    upd = t * np.float32(growth)
    if save:
        d = os.path.join(out_dir or P.IO_DIR, "Boxes")
        os.makedirs(d, exist_ok=True)
        np.save(os.path.join(d, updated_name(z, upd.shape[0], L)), upd.cpu().numpy())
    return upd


def velocity_name(axis, HII_DIM, BOX_LEN):
    return "updated_v{0}overddot_{1:d}_{2:.0f}Mpc".format(axis, int(HII_DIM), float(BOX_LEN))


def velocities_at(z, deltax, HII_DIM=None, box_len=None, ddt_over_d=None, save=False, out_dir=None):
    """Velocity boxes at z from the z = 0 linear density (perturb_field.py:
    111-126): v = (dD/dt)/D(z) * psi, psi~ = i k / k^2 delta~, top-hat filtered
    at L_FACTOR * BOX_LEN / HII_DIM, inverse transformed and subsampled to
    HII_DIM (the initialize velocity path with the growth-rate factor).
    deltax: DIM^3 CUDA tensor.  ddt_over_d: (dD/dt)/D in s^-1, default the
    Planck13 value (cosmo_files.planck).  Returns (vx, vy, vz) in the array-axis
    order of initialize.init's velocities."""
    _lib.require_gpu()
    L = float(P.BOX_LEN if box_len is None else box_len)
    HII = int(P.HII_DIM if HII_DIM is None else HII_DIM)
    DIM = int(deltax.shape[0])
    if DIM % HII:
        raise ValueError(f"HII_DIM={HII} must divide DIM={DIM}")
    f = planck.ddt_growth_over_growth(float(z)) if ddt_over_d is None else float(ddt_over_d)
    spec = gpu.r2c(deltax)  # DIM^3 times the c2r-ready spectrum
    R = initialize.smoothing_radius(L, HII)
    work = torch.empty_like(spec)
    large = torch.empty((DIM, DIM, DIM), dtype=torch.float32, device=deltax.device)
    scale = np.float32(f / float(DIM) ** 3)
    vs = []
    for comp in range(3):
        gpu.velocity(spec, L, comp, filter_type=0, R=R, out=work)
        gpu.c2r(work, out=large, preserve_input=False)
        v = gpu.subsample(large, HII)
        v.mul_(scale)
        vs.append(v)
    if save:
        d = os.path.join(out_dir or P.IO_DIR, "Boxes")
        os.makedirs(d, exist_ok=True)
        for axis, v in zip("xyz", vs):
            np.save(os.path.join(d, velocity_name(axis, HII, L)), v.cpu().numpy())
    return tuple(vs)


def run(z, growth, growth_init, DIM=None, HII_DIM=None, BOX_LEN=None, scheme="cic", linear=None, io_dir=None):
    """perturb_field.py:130-144: reads the initialize outputs from Boxes/."""
    DIM = int(P.DIM if DIM is None else DIM)
    HII_DIM = int(P.HII_DIM if HII_DIM is None else HII_DIM)
    L = float(P.BOX_LEN if BOX_LEN is None else BOX_LEN)
    d = os.path.join(io_dir or P.IO_DIR, "Boxes")
    linear = P.EVOLVE_DENSITY_LINEARLY if linear is None else linear
    if linear:
        x = np.load(os.path.join(d, "smoothed_deltax_z0.00_{0:d}_{1:.0f}Mpc.npy".format(HII_DIM, L)))
        return evolve_linear(z, x, growth, L, save=True, out_dir=io_dir)
    x = torch.from_numpy(np.load(os.path.join(d, "deltax_z0.00_{0:d}_{1:.0f}Mpc.npy".format(DIM, L)))).cuda()
    v = tuple(torch.from_numpy(np.load(os.path.join(d, "v{0}overddot_{1:d}_{2:.0f}Mpc.npy".format(m, HII_DIM, L))))
              .cuda() for m in "xyz")
    return evolve_zeldovich(z, x, v, growth, growth_init, L, scheme, save=True, out_dir=io_dir)
