"""Zel'dovich mass move (perturb_field stage) on the CPU.

TEST INFRASTRUCTURE ONLY (see ``oracle/__init__.py``).

``ngp_targets_f32`` replays Programs/perturb_field.cu:90-127 in float32 (particle
at the cell centre, displacement from the HII-resolution field, scale to the
HII grid, wrap, truncate), ``move_mass_ngp`` adds the deposit of :132 and the
renormalisation of perturb_field.py:104-105.  ``move_mass_cic`` is the
cloud-in-cell deposit of the same particles (no reference counterpart), in FP64.
"""
import numpy as np


def _f32(x):
    return np.asarray(x, dtype=np.float32)


def displacement_scale(growth, growth_init, box_len):
    """(D(z) - D_init) / BOX_LEN in float32, perturb_field.py:93-95."""
    return np.float32(_f32(np.float32(growth) - np.float32(growth_init)) / np.float32(box_len))


def particle_positions_f32(DIM, vx, vy, vz, disp_scale):
    """float32 HII-grid positions (xf, yf, zf) of all DIM^3 particles, after the
    wrap loops of perturb_field.cu:113-118.  Arrays are (DIM, DIM, DIM) with
    axis 0 = i (x), axis 2 = k (z), as C_INDEX."""
    sw = vx.shape[0]
    swf = np.float32(sw)
    pf = float(np.float32(DIM) / np.float32(sw))
    idx = np.arange(DIM)
    base = _f32((idx + 0.5) / float(DIM))                  # :102-104 (double divide)
    hii = np.floor(idx / pf).astype(np.int64)              # :105-107
    I, J, K = np.ix_(hii, hii, hii)
    out = []
    for comp, v in enumerate((vx, vy, vz)):
        vs = _f32(_f32(v) * np.float32(disp_scale))[I, J, K]
        shape = [1, 1, 1]
        shape[comp] = DIM
        f = _f32(base.reshape(shape) + vs)                  # :108-110
        f = _f32(f * swf)                                   # :111
        while True:                                         # :113-118
            hi = f >= swf
            lo = f < 0
            if not (hi.any() or lo.any()):
                break
            f = np.where(hi, _f32(f - swf), f)
            f = np.where(lo, _f32(f + swf), f)
        out.append(f)
    return out


def ngp_targets_f32(DIM, vx, vy, vz, disp_scale):
    """Integer NGP target cells (xi, yi, zi), perturb_field.cu:119-127."""
    sw = vx.shape[0]
    res = []
    for f in particle_positions_f32(DIM, vx, vy, vz, disp_scale):
        i = np.trunc(f).astype(np.int64)
        i = np.where(i >= sw, i - sw, i)
        i = np.where(i < 0, i + sw, i)
        res.append(i)
    return res


def particle_mass_f32(delta, growth_init):
    """float(1.0 + init_growth*delta) as passed to atomicAdd, perturb_field.cu:132."""
    return _f32(1.0 + _f32(np.float32(growth_init) * _f32(delta)).astype(np.float64))


def renormalise(acc, DIM, sw):
    """updated / MASS_FACTOR - 1, perturb_field.py:104-105 (float32)."""
    mass_factor = np.float32((np.float32(DIM) / np.float32(sw)) ** 3)
    return _f32(_f32(_f32(acc) / mass_factor) - np.float32(1.0))


def move_mass_ngp(delta, vx, vy, vz, growth, growth_init, box_len):
    DIM = delta.shape[0]
    sw = vx.shape[0]
    s = displacement_scale(growth, growth_init, box_len)
    xi, yi, zi = ngp_targets_f32(DIM, vx, vy, vz, s)
    m = particle_mass_f32(delta, growth_init).astype(np.float64)
    flat = (xi * sw + yi) * sw + zi
    acc = np.bincount(flat.ravel(), weights=m.ravel(), minlength=sw ** 3).reshape(sw, sw, sw)
    return renormalise(acc, DIM, sw)


def move_mass_cic(delta, vx, vy, vz, growth, growth_init, box_len):
    """Cloud-in-cell deposit: cell c covers [c, c+1) with centre c + 0.5, so a
    particle at f shares its mass between floor(f - 0.5) and the next cell with
    weights (1 - t, t), t = f - 0.5 - floor(f - 0.5), periodic."""
    DIM = delta.shape[0]
    sw = vx.shape[0]
    s = displacement_scale(growth, growth_init, box_len)
    pos = particle_positions_f32(DIM, vx, vy, vz, s)
    m = particle_mass_f32(delta, growth_init).astype(np.float64).ravel()
    lo, t = [], []
    for f in pos:
        x = f.astype(np.float64).ravel() - 0.5
        i0 = np.floor(x)
        lo.append(i0.astype(np.int64))
        t.append(x - i0)
    acc = np.zeros(sw ** 3)
    for dx in (0, 1):
        wx = t[0] if dx else 1 - t[0]
        ix = (lo[0] + dx) % sw
        for dy in (0, 1):
            wy = t[1] if dy else 1 - t[1]
            iy = (lo[1] + dy) % sw
            for dz in (0, 1):
                wz = t[2] if dz else 1 - t[2]
                iz = (lo[2] + dz) % sw
                acc += np.bincount((ix * sw + iy) * sw + iz, weights=m * wx * wy * wz, minlength=sw ** 3)
    return renormalise(acc.reshape(sw, sw, sw), DIM, sw)


def evolve_linear(delta, growth):
    """perturb_field.py:39-50: delta * D(z) in float32."""
    return _f32(_f32(delta) * np.float32(growth))
