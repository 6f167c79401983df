"""Excursion-set ionization (find_HII_bubbles) restated on the CPU.

TEST INFRASTRUCTURE ONLY (see ``oracle/__init__.py``).

Follows Programs/find_bubbles.cu:12-158 (filter, f_coll, update, sphere
paint, final paint) and Programs/find_HII_bubbles.py:53-193 (control flow).
The FFTs are NumPy's FP64 real transforms; every other step of the ``_f32``
path replays the reference kernels' float32 operation order, so that integer
decisions (the k-space cut-off, the threshold away from its 1e-5 band) are
reproduced exactly.
"""
import numpy as np
from scipy.special import erfc as _erfc

from .sigma import mode_index

E_F32 = np.float32(2.7182818284)          # find_bubbles.cu:3
KTOPHAT_FACTOR = 0.413566994              # find_bubbles.cu:34
GAUSS_FACTOR = 0.643                      # find_bubbles.cu:39


def _f32(x):
    return np.asarray(x, dtype=np.float32)


def delta_k_f32(box_len):
    """DELTA_K = float32(2 pi / float32(BOX_LEN)), Parameter_files/INIT_PARAMS.py:9-10."""
    return np.float32(2.0 * np.pi / float(np.float32(box_len)))


def kmag_half_f32(D, dk, fastest_first=True):
    """|k| in float32 on the r2c half grid (D, D, D//2+1), axis 2 fastest.

    find_bubbles.cu:22-26 builds k_c = n_c * DELTAK and sums the squares with the
    fastest axis first (INDEX layout); initialize.cu:20-24 sums the slowest axis
    first (C_INDEX layout).  ``fastest_first`` picks the order.
    """
    n_full = mode_index(D).astype(np.float32)
    n_half = np.arange(D // 2 + 1).astype(np.float32)
    k0 = (n_full * np.float32(dk)).astype(np.float32)[:, None, None]
    k1 = (n_full * np.float32(dk)).astype(np.float32)[None, :, None]
    k2 = (n_half * np.float32(dk)).astype(np.float32)[None, None, :]
    sq0, sq1, sq2 = _f32(k0 * k0), _f32(k1 * k1), _f32(k2 * k2)
    if fastest_first:
        s = _f32(_f32(sq2 + sq1) + sq0)
    else:
        s = _f32(_f32(sq0 + sq1) + sq2)
    return np.sqrt(s).astype(np.float32)


def window_f32(kmag, R, filter_type, ref_fallthrough=True):
    """Window multiplier per mode, replaying find_bubbles.cu:27-41 in float32.

    Returns (w0, cut, w2): the spectrum value becomes 0 where ``cut`` is true and
    ((value * w0) * w2) otherwise, each product rounded to float32.
    ref_fallthrough=True reproduces the switch without breaks (type 0 runs
    cases 0,1,2; type 1 runs 1,2; type 2 runs 2).  False gives the textbook
    windows: W_TH(kR), the k-space cut at kR*0.413566994 > 1, or the Gaussian.
    """
    kmag = _f32(kmag)
    kR = _f32(kmag * np.float32(R))
    w0 = np.ones_like(kR)
    cut = np.zeros(kR.shape, dtype=bool)
    w2 = np.ones_like(kR)
    if filter_type == 0:
        big = kR.astype(np.float64) > 1e-4
        kRs = np.where(big, kR, np.float32(1.0)).astype(np.float64)
        s = _f32(np.sin(kRs)); c = _f32(np.cos(kRs))
        p3 = _f32(kRs * kRs * kRs); p2 = _f32(kRs * kRs)
        inner = _f32(_f32(s / p3) - _f32(c / p2))
        w0 = np.where(big, _f32(3.0 * inner.astype(np.float64)), np.float32(1.0)).astype(np.float32)
    if not ref_fallthrough:
        if filter_type == 1:
            cut = _f32(kR.astype(np.float64) * KTOPHAT_FACTOR) > np.float32(1.0)
        elif filter_type == 2:
            kg = _f32(kR.astype(np.float64) * GAUSS_FACTOR)
            w2 = np.power(E_F32, _f32(-_f32(kg * kg) / np.float32(2.0))).astype(np.float32)
        return w0, cut, w2
    if filter_type in (0, 1):
        kR = _f32(kR.astype(np.float64) * KTOPHAT_FACTOR)
        cut = kR > np.float32(1.0)
    kR = _f32(kR.astype(np.float64) * GAUSS_FACTOR)
    w2 = np.power(E_F32, _f32(-_f32(kR * kR) / np.float32(2.0))).astype(np.float32)
    return w0, cut, w2


def window_f64(kmag, R, filter_type, ref_fallthrough=True):
    """The same windows as one FP64 multiplier (0 where cut)."""
    kR = np.asarray(kmag, dtype=np.float64) * float(R)
    w = np.ones_like(kR)
    if filter_type == 0:
        big = kR > 1e-4
        ks = np.where(big, kR, 1.0)
        w = np.where(big, 3.0 * (np.sin(ks) / ks ** 3 - np.cos(ks) / ks ** 2), 1.0)
    if not ref_fallthrough:
        if filter_type == 1:
            w = np.where(kR * KTOPHAT_FACTOR > 1.0, 0.0, 1.0)
        elif filter_type == 2:
            w = np.exp(-(kR * GAUSS_FACTOR) ** 2 / 2.0)
        return w
    if filter_type in (0, 1):
        kR = kR * KTOPHAT_FACTOR
        w = np.where(kR > 1.0, 0.0, w)
    kR = kR * GAUSS_FACTOR
    return w * np.exp(-kR * kR / 2.0)


def apply_window_f32(spec, w0, cut, w2):
    """value -> 0 if cut else (value*w0)*w2, products on the float32 spectrum."""
    spec = spec.astype(np.complex64)
    re = _f32(_f32(spec.real * w0) * w2)
    im = _f32(_f32(spec.imag * w0) * w2)
    re[cut] = 0
    im[cut] = 0
    return (re + 1j * im).astype(np.complex64)


def erfc_f32(x):
    """erfcf on float32 input: FP64 erfc rounded to float32."""
    return _f32(_erfc(np.asarray(x, dtype=np.float64)))


def fcoll_f32(smoothed, deltac, denom):
    """find_bubbles.cu:53-57: erfcf((deltac - delta0)/denom), no clamp."""
    return erfc_f32(_f32(_f32(np.float32(deltac) - _f32(smoothed)) / np.float32(denom)))


def nonlinear_input_f32(d, growth_inv):
    """Linear extrapolation of a nonlinear density box to z = 0,
    find_HII_bubbles.py:58,84-85: deltax *= float32(deltac/1.686)."""
    return _f32(_f32(d) * np.float32(growth_inv))


def nonlinear_input_f64(d, growth_inv):
    """FP64 counterpart of nonlinear_input_f32."""
    return np.asarray(d, dtype=np.float64) * float(growth_inv)


def box_mean_scale_f32(fc, fc_mean):
    """Normalisation factor of find_HII_bubbles.py:158-161,171-174:
    float32(fc_mean / <f_coll>) with the box mean <f_coll> = sum/N accumulated
    in FP64 (the reference sums in float32 on the GPU; the build keeps FP64)."""
    mean = np.float64(np.sum(np.asarray(fc, dtype=np.float64))) / np.asarray(fc).size
    with np.errstate(divide="ignore", invalid="ignore", over="ignore"):
        # a box mean of 0 gives an infinite scale (IEEE), as in the reference
        return np.float32(np.float64(np.float32(fc_mean)) / mean)


def box_mean_scale_f64(fc, fc_mean):
    """FP64 counterpart of box_mean_scale_f32."""
    with np.errstate(divide="ignore", invalid="ignore"):
        return np.float64(fc_mean) / np.float64(np.mean(fc))


def paint_spheres(ionized, centers, R, periodic=False):
    """Sphere paint of find_bubbles.cu:83-101 for a boolean ``centers`` mask.

    Every cell whose integer squared distance to some centre is below
    float32(R*R) is set to 1.  The reference clips the bounding box at the
    edges (non-periodic, :87-92); ``periodic`` wraps instead.  Computed with an
    exact Euclidean distance transform (scipy.ndimage).
    """
    from scipy import ndimage
    if not centers.any():
        return ionized
    R2 = np.float32(np.float32(R) * np.float32(R))
    D = centers.shape[0]
    if periodic:
        pad = int(np.ceil(float(R))) + 1
        pad = min(pad, D)
        tiled = np.pad(centers, pad, mode="wrap")
        dist = ndimage.distance_transform_edt(~tiled, return_distances=True)
        d2 = np.rint(dist ** 2)[pad:pad + D, pad:pad + D, pad:pad + D]
    else:
        dist = ndimage.distance_transform_edt(~centers, return_distances=True)
        d2 = np.rint(dist ** 2)
    ionized[d2.astype(np.float32) < R2] = np.float32(1.0)
    return ionized


def find_bubbles_f32(delta, box_len, radii, denoms, zeta, deltac, LE=True,
                     growth_inv=1.0, fc_mean=None, fil=1, ref_fallthrough=True,
                     update=0, sphere_periodic=False, sphere_r_scale=1.0,
                     return_trace=False):
    """FP32-faithful restatement of conv_bubbles (find_HII_bubbles.py:17-193).

    delta: float32 (D,D,D).  radii/denoms: float32 lists from the schedule.
    LE=False multiplies delta by ``growth_inv`` (= deltac/1.686, :58,84-85) and
    normalises f_coll each radius by fc_mean / <f_coll> (:158-161,171-174) with
    <f_coll> accumulated in FP64.  update=0 is the centre-pixel rule, update=1
    the sphere paint (R times ``sphere_r_scale`` cells; 1.0 replays the reference's
    R-in-Mpc-as-cells).
    Returns the float32 ionized box (0, 1 or partial zeta*f_coll).
    """
    d = _f32(delta)
    D = d.shape[0]
    if not LE:
        d = nonlinear_input_f32(d, growth_inv)
    spec = np.fft.rfftn(d.astype(np.float64)).astype(np.complex64)
    kmag = kmag_half_f32(D, delta_k_f32(box_len), fastest_first=True)
    ionized = np.zeros((D, D, D), dtype=np.float32)
    inv_zeta = 1.0 / float(zeta)
    trace = []
    nR = len(radii)
    for r, (R, denom) in enumerate(zip(radii, denoms)):
        final = r == nR - 1
        w0, cut, w2 = window_f32(kmag, R, fil, ref_fallthrough)
        filt = apply_window_f32(spec, w0, cut, w2)
        sm = _f32(np.fft.irfftn(filt.astype(np.complex128), s=(D, D, D), axes=(0, 1, 2)))
        fc = fcoll_f32(sm, deltac, denom)
        scale = None
        if not LE:
            scale = box_mean_scale_f32(fc, fc_mean)
            with np.errstate(invalid="ignore"):
                fc = _f32(fc * scale)
        hit = fc.astype(np.float64) >= inv_zeta
        if return_trace:
            trace.append(dict(R=R, smoothed=sm, fcoll=fc, scale=scale))
        if final:
            todo = ionized != np.float32(1.0)
            ionized[todo & hit] = 1.0
            part = todo & ~hit
            ionized[part] = _f32(fc[part].astype(np.float64) * float(zeta))
        elif update == 0:
            ionized[hit] = 1.0
        else:
            paint_spheres(ionized, hit, np.float32(float(R) * sphere_r_scale), sphere_periodic)
    if return_trace:
        return ionized, trace
    return ionized


def find_bubbles_f64(delta, box_len, radii, denoms, zeta, deltac, LE=True,
                     growth_inv=1.0, fc_mean=None, fil=1, ref_fallthrough=True, fft_workers=None):
    """FP64 version of the centre-pixel loop; also returns, per cell, the
    minimum over radii of |zeta*f_coll - 1| (the threshold margin).
    fft_workers: run the transforms with scipy.fft on that many threads and
    the per-cell work on as many threads over slabs of axis 0 (same results up
    to FP64 rounding; for large boxes)."""
    import scipy.fft as sfft
    from concurrent.futures import ThreadPoolExecutor
    workers = int(fft_workers) if fft_workers else 1
    if fft_workers:
        rfftn = lambda a: sfft.rfftn(a, workers=workers)  # noqa: E731
        irfftn = lambda a, s: sfft.irfftn(a, s=s, axes=(0, 1, 2), workers=workers)  # noqa: E731
    else:
        rfftn = np.fft.rfftn
        irfftn = lambda a, s: np.fft.irfftn(a, s=s, axes=(0, 1, 2))  # noqa: E731
    d = np.asarray(delta, dtype=np.float64)
    D = d.shape[0]
    if not LE:
        d = nonlinear_input_f64(d, growth_inv)
    spec = rfftn(d)
    dk = 2.0 * np.pi / float(box_len)
    n0 = mode_index(D).astype(np.float64)
    n2 = np.arange(D // 2 + 1, dtype=np.float64)
    kmag = dk * np.sqrt(n0[:, None, None] ** 2 + n0[None, :, None] ** 2 + n2[None, None, :] ** 2)
    ionized = np.zeros((D, D, D))
    margin = np.full((D, D, D), np.inf)
    fc = np.empty((D, D, D))
    nR = len(radii)
    step = -(-D // workers)
    slabs = [slice(i, min(i + step, D)) for i in range(0, D, step)]
    ex = ThreadPoolExecutor(workers) if workers > 1 else None

    def each(fn):
        return list(ex.map(fn, slabs)) if ex else [fn(sl) for sl in slabs]

    try:
        for r, (R, denom) in enumerate(zip(radii, denoms)):
            filt = np.empty_like(spec)

            def window(sl, R=R):
                filt[sl] = spec[sl] * window_f64(kmag[sl], float(R), fil, ref_fallthrough)

            each(window)
            sm = irfftn(filt, (D, D, D))
            del filt

            def fcoll(sl, sm=sm, denom=denom):
                fc[sl] = _erfc((float(deltac) - sm[sl]) / float(denom))
                return float(np.sum(fc[sl]))

            total = sum(each(fcoll))
            scale = None
            if not LE:
                with np.errstate(divide="ignore", invalid="ignore"):
                    scale = np.float64(fc_mean) / (np.float64(total) / fc.size)
            final = r == nR - 1

            def update(sl, scale=scale, final=final):
                f = fc[sl] if scale is None else fc[sl] * scale
                np.minimum(margin[sl], np.abs(zeta * f - 1.0), out=margin[sl])
                hit = f >= 1.0 / zeta
                ion = ionized[sl]
                if final:
                    todo = ion != 1.0
                    ion[todo & hit] = 1.0
                    part = todo & ~hit
                    ion[part] = f[part] * zeta
                else:
                    ion[hit] = 1.0

            with np.errstate(invalid="ignore"):
                each(update)
    finally:
        if ex:
            ex.shutdown()
    return ionized, margin
