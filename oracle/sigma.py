"""Mass variance S(R), synthetic power spectrum and the radius schedule.

TEST INFRASTRUCTURE ONLY (see ``oracle/__init__.py``).

* ``sig0_trapz`` restates cosmo_files/sigmas.py:23-24,54-66,76-78,99-102,
  176-191,224-232: S(R) = trapz over ln k of Delta^2(k) W_TH(kR)^2 on 1000
  points between limits found on a 0.1-spaced ln k grid where the integrand
  exceeds 1e-6 of its maximum.
* ``synthetic_*`` implement the configs' spectrum P(k) = A k^-2.5 exp(-(k/k_nyq)^2)
  with A fixed by (1/V) sum_{k != 0} P(k) = sigma_cell^2 over the full D^3 grid.
* ``radius_schedule`` restates Programs/find_HII_bubbles.py:53-54,92,113-125,189
  with NumPy-1.x scalar promotion written out explicitly (float32 R, float64
  R/1.1), and ``denominators`` restates :126-128.
"""
import numpy as np

L_FACTOR = np.float32(0.620350491)  # Parameter_files/INIT_PARAMS.py:31
R_BUBBLE_MAX = 30.0                 # INIT_PARAMS.py:21
R_BUBBLE_MIN = 0.5                  # INIT_PARAMS.py:22
DELTA_R_FACTOR = 1.1                # INIT_PARAMS.py:24


def tophat_window(y):
    """Real-space top-hat window in k-space, cosmo_files/sigmas.py:24."""
    y = np.asarray(y, dtype=np.float64)
    return 3.0 * (np.sin(y) - y * np.cos(y)) / y ** 3


def tophat_window_stable(y):
    """W_TH with the small-argument series 1 - y^2/10 + y^4/280 below |y| < 1e-2
    (the closed form loses all digits to cancellation there)."""
    y = np.asarray(y, dtype=np.float64)
    small = np.abs(y) < 1.0e-2
    ys = np.where(small, 1.0, y)
    big = 3.0 * (np.sin(ys) - ys * np.cos(ys)) / ys ** 3
    y2 = y * y
    return np.where(small, 1.0 - y2 / 10.0 + y2 * y2 / 280.0, big)


def _klims(vals, logk, factor=1.0e-6):
    """Integration limits, cosmo_files/sigmas.py:54-66."""
    a = np.abs(vals)
    mask = a > a.max() * factor
    return logk[mask].min(), logk[mask].max()


def sig0_trapz(R, del2_fn, nint=1000):
    """S(R) by the reference's trapz rule, cosmo_files/sigmas.py:176-191."""
    logk = np.arange(-20.0, 20.0, 0.1)
    k = np.exp(logk)
    lo, hi = _klims(del2_fn(k) * tophat_window(R * k) ** 2, logk)
    lk = np.linspace(lo, hi, nint)
    kk = np.exp(lk)
    return float(np.trapezoid(del2_fn(kk) * tophat_window(R * kk) ** 2, lk))


def table_del2(k_tab, del2_tab):
    """Linear interpolation of a tabulated Delta^2(k), as interp1d at sigmas.py:52."""
    k_tab = np.asarray(k_tab, dtype=np.float64)
    d_tab = np.asarray(del2_tab, dtype=np.float64)
    return lambda k: np.interp(k, k_tab, d_tab)


# ---------------------------------------------------------------- synthetic P(k)

def mode_index(D):
    """Signed mode number per FFT index, with index D/2 mapped to +D/2
    (find_bubbles.cu:22-24: k = (i > w/2) ? (i-w) dk : i dk)."""
    n = np.arange(D)
    return np.where(n > D // 2, n - D, n)


def nsq_histogram_bruteforce(D):
    """Multiplicity of each integer |n|^2 over the full D^3 grid, by direct bincount."""
    n = mode_index(D) ** 2
    out = np.zeros(3 * (D // 2) ** 2 + 1, dtype=np.int64)
    plane = (n[:, None] + n[None, :]).ravel()
    for nz in n:
        out += np.bincount(plane + nz, minlength=out.size)
    return out


def synthetic_pk_shape(k, k_nyq):
    """Unnormalised P(k) = k^-2.5 exp(-(k/k_nyq)^2), P(0) = 0."""
    k = np.asarray(k, dtype=np.float64)
    out = np.zeros_like(k)
    nz = k > 0
    out[nz] = k[nz] ** -2.5 * np.exp(-(k[nz] / k_nyq) ** 2)
    return out


def synthetic_amplitude(D, box_len, sigma_cell2=0.36):
    """A such that (1/V) sum_{k != 0} A * shape(k) = sigma_cell2 on the D^3 grid."""
    dk = 2.0 * np.pi / box_len
    k_nyq = np.pi * D / box_len
    h = nsq_histogram_bruteforce(D)
    m = np.arange(h.size)
    s = np.sum(h[1:] * synthetic_pk_shape(dk * np.sqrt(m[1:]), k_nyq))
    V = float(box_len) ** 3
    return sigma_cell2 * V / s


def synthetic_pk(k, D, box_len, sigma_cell2=0.36):
    A = synthetic_amplitude(D, box_len, sigma_cell2)
    return A * synthetic_pk_shape(k, np.pi * D / box_len)


def sig0_discrete(R, D, box_len, sigma_cell2=0.36):
    """S(R) = (1/V) sum_{k != 0} P(k) W_TH(kR)^2 over the box's own modes (FP64)."""
    dk = 2.0 * np.pi / box_len
    k_nyq = np.pi * D / box_len
    A = synthetic_amplitude(D, box_len, sigma_cell2)
    h = nsq_histogram_bruteforce(D)
    m = np.arange(1, h.size)
    k = dk * np.sqrt(m)
    V = float(box_len) ** 3
    return float(np.sum(h[1:] * A * synthetic_pk_shape(k, k_nyq) * tophat_window_stable(k * R) ** 2) / V)


def synthetic_del2_fn(D, box_len, sigma_cell2=0.36):
    """Delta^2(k) = k^3 P(k) / (2 pi^2) for the synthetic spectrum (continuum form)."""
    A = synthetic_amplitude(D, box_len, sigma_cell2)
    k_nyq = np.pi * D / box_len
    return lambda k: np.asarray(k, dtype=np.float64) ** 3 * A * synthetic_pk_shape(k, k_nyq) / (2 * np.pi ** 2)


# ---------------------------------------------------------------- radius schedule

def radius_schedule(box_len, dim, py2_int_scale=False, scale=None):
    """float32 radii from R_BUBBLE_MAX down, and the cell size used by the stop rule.

    find_HII_bubbles.py:52-54 (a caller's scale, else scale = BoxSize/HIIdim,
    integer division under Python 2 when both are ints), :92 (R =
    R_BUBBLE_MAX), :123-125 (final_step test done in float64 before the
    float32 cast), :189 (R = R/1.1 in float64).
    """
    if scale is None:
        scale = (int(box_len) // int(dim)) if py2_int_scale else float(box_len) / float(dim)
    stop = float(L_FACTOR) * scale
    R = float(R_BUBBLE_MAX)
    radii = []
    final = False
    while not final:
        if (R / DELTA_R_FACTOR) <= stop or (R / DELTA_R_FACTOR) <= R_BUBBLE_MIN:
            final = True
        R32 = np.float32(R)
        radii.append(R32)
        R = float(R32) / DELTA_R_FACTOR
    return np.array(radii, dtype=np.float32), scale


def denominators(radii, sigma_min2, sig0_fn):
    """denom = float32(sqrt(2 (smin - float32(S(R))))), find_HII_bubbles.py:126-128."""
    out = []
    for R in radii:
        S0 = np.float32(sig0_fn(float(R)))
        arg = 2.0 * (float(sigma_min2) - float(S0))
        out.append(np.float32(np.sqrt(arg)) if arg >= 0 else np.float32(np.nan))
    return np.array(out, dtype=np.float32)
