"""Density field at redshift z (drop-in for Programs/perturb_field.py).

The entry points keep the reference's calls: ``run(z)``,
``evolve_zeldovich(z, deltax)`` and ``evolve_linear(z, deltax)``
(perturb_field.py:39,65,130-144), with the same defaults: growth factors
D(z) and D(INITIAL_REDSHIFT) of the Planck13 cosmology (cosmo_files.planck.
fgrowth, which restates cosmolopy's fgrowth), displacement boxes read from
Boxes/ (perturb_field.py:87-89), the updated box written under the reference's
name (:50, :108) and, for the Zel'dovich move, the velocity boxes at z
(:111-126).  Keyword overrides give the growth factors, the displacement boxes
and the box size directly, and switch the file output off.

``evolve_zeldovich`` follows perturb_field.py:65-108: the displacement boxes
are scaled by (D(z) - D_init)/BOX_LEN, the DIM^3 particles move and deposit
(1 + D_init delta) on the HII grid, and the result is divided by MASS_FACTOR
minus one.  ``scheme='ngp'`` (default) is the reference's truncation; ``'cic'``
deposits cloud-in-cell.  ``evolve_linear`` is perturb_field.py:39-50.
"""
import os

import numpy as np
import torch

from .. import gpu, _lib
from ..Parameter_files import INIT_PARAMS as P
from ..cosmo_files import planck
from ..tocmfastpy import boxio
from . import initialize


def displacement_scale(growth, growth_init, box_len):
    """(fgrowth - primordial_fgrowth)/BOX_LEN in float32 (perturb_field.py:93-95)."""
    return np.float32(np.float32(np.float32(growth) - np.float32(growth_init)) / np.float32(box_len))


def updated_name(z, HII_DIM, BOX_LEN):
    return "updated_smoothed_deltax_z{0:.2f}_{1:d}_{2:.0f}Mpc".format(float(z), int(HII_DIM), float(BOX_LEN))


def _boxes_dir(out_dir):
    return os.path.join(out_dir or P.IO_DIR, "Boxes")


def _device_box(x):
    if isinstance(x, torch.Tensor):
        return x.to(torch.device("cuda"), dtype=torch.float32).contiguous()
    return torch.from_numpy(np.ascontiguousarray(x, dtype=np.float32)).cuda()


def _load_box(path_npy):
    """A box saved as .npy, or else the raw 21cmFAST file of the same name
    without the extension (perturb_field.py:134-143)."""
    if os.path.exists(path_npy):
        return np.load(path_npy)
    return boxio.readbox(os.path.splitext(path_npy)[0]).box_data


def load_velocities(HII_DIM=None, BOX_LEN=None, out_dir=None):
    """The initialize displacement boxes v{x,y,z}overddot (perturb_field.py:87-89)."""
    HII = int(P.HII_DIM if HII_DIM is None else HII_DIM)
    L = float(P.BOX_LEN if BOX_LEN is None else BOX_LEN)
    names = initialize.box_names(HII, HII, L)["v"]
    return tuple(_device_box(_load_box(os.path.join(_boxes_dir(out_dir), n + ".npy"))) for n in names)


def evolve_zeldovich(z, deltax, v=None, growth=None, growth_init=None, box_len=None, scheme="ngp", save=True,
                     out_dir=None, velocities=None, HII_DIM=None):
    """Zel'dovich mass move of the DIM^3 density ``deltax`` (NumPy or CUDA
    tensor) to z on the HII grid; returns the updated density (CUDA tensor).

    v: (vx, vy, vz) HII_DIM^3 displacement boxes, default read from Boxes/.
    growth, growth_init: D(z) and D(INITIAL_REDSHIFT), default Planck13.
    save: write updated_smoothed_deltax_z..._Mpc.npy (and, unless velocities
    is False, the velocity boxes at z) as the reference does."""
    _lib.require_gpu()
    L = float(P.BOX_LEN if box_len is None else box_len)
    if growth is None:
        growth = np.float32(planck.fgrowth(float(z)))
    if growth_init is None:
        growth_init = np.float32(planck.fgrowth(float(P.INITIAL_REDSHIFT)))
    d = _device_box(deltax)
    if v is None:
        v = load_velocities(HII_DIM, L, out_dir)
    s = displacement_scale(growth, growth_init, L)
    upd = gpu.move_mass(d, v[0], v[1], v[2], s, np.float32(growth_init), scheme)
    if save:
        bd = _boxes_dir(out_dir)
        os.makedirs(bd, exist_ok=True)
        np.save(os.path.join(bd, updated_name(z, upd.shape[0], L)), upd.cpu().numpy())
        if velocities is None or velocities:
            velocities_at(z, d, HII_DIM=upd.shape[0], box_len=L, save=True, out_dir=out_dir)
    return upd


def evolve_linear(z, deltax, growth=None, box_len=None, save=True, out_dir=None):
    """delta * D(z) (perturb_field.py:39-50); growth defaults to the Planck13 D(z)."""
    L = float(P.BOX_LEN if box_len is None else box_len)
    if growth is None:
        growth = np.float32(planck.fgrowth(float(z)))
    upd = _device_box(deltax) * np.float32(growth)
    if save:
        bd = _boxes_dir(out_dir)
        os.makedirs(bd, exist_ok=True)
        np.save(os.path.join(bd, updated_name(z, upd.shape[0], L)), upd.cpu().numpy())
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


def run(z, growth=None, growth_init=None, DIM=None, HII_DIM=None, BOX_LEN=None, scheme="ngp", linear=None,
        io_dir=None):
    """perturb_field.py:130-144: reads the initialize outputs from Boxes/
    (.npy, else the raw box of the same name) and writes the updated box."""
    DIM = int(P.DIM if DIM is None else DIM)
    HII_DIM = int(P.HII_DIM if HII_DIM is None else HII_DIM)
    L = float(P.BOX_LEN if BOX_LEN is None else BOX_LEN)
    d = _boxes_dir(io_dir)
    names = initialize.box_names(DIM, HII_DIM, L)
    linear = P.EVOLVE_DENSITY_LINEARLY if linear is None else linear
    if linear:
        x = _load_box(os.path.join(d, names["smoothed"] + ".npy"))
        return evolve_linear(z, x, growth, L, save=True, out_dir=io_dir)
    x = _load_box(os.path.join(d, names["deltax"] + ".npy"))
    return evolve_zeldovich(z, x, None, growth, growth_init, L, scheme, save=True, out_dir=io_dir,
                            HII_DIM=HII_DIM)
