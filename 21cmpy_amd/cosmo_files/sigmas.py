"""Mass variance S(R) = sigma^2(R) for the excursion-set barrier.

The reference tabulates Delta^2(k) for Planck13 with cosmolopy and evaluates
S(R) by the trapezoid rule over ln k (cosmo_files/sigmas.py:44-52,176-191,
224-232).  cosmolopy is not available, so this module works with an explicit
spectrum object.  ``PowerLawSpectrum`` is the synthetic spectrum of the
benchmark configurations, P(k) = amp k^n exp(-(k/k_cut)^2), normalised so that
the box's own modes carry a given per-cell variance:
(1/V) sum_{k != 0} P(k) = sigma_cell^2 over the D^3 grid.

``sig0`` offers two definitions of S(R):
* ``"discrete"`` (default): (1/V) sum over the box's modes of P(k) W_TH(kR)^2,
  consistent with the normalisation above;
* ``"trapz"``: the reference's continuum rule, 1000-point trapezoid over ln k
  between the limits where Delta^2 W^2 exceeds 1e-6 of its maximum on a
  0.1-spaced ln k grid (sigmas.py:54-66,99-102,176-191).
"""
import functools

import numpy as np


def tophat_window(y):
    """W_TH(y) = 3 (sin y - y cos y) / y^3, with its series below |y| < 1e-2."""
    y = np.asarray(y, dtype=np.float64)
    small = np.abs(y) < 1.0e-2
    ys = np.where(small, 1.0, y)
    closed = 3.0 * (np.sin(ys) - ys * np.cos(ys)) / ys ** 3
    y2 = y * y
    return np.where(small, 1.0 - y2 / 10.0 + y2 * y2 / 280.0, closed)


def tophat_window_closed(y):
    """The closed form exactly as the reference evaluates it (sigmas.py:24)."""
    y = np.asarray(y, dtype=np.float64)
    return 3.0 * (np.sin(y) - y * np.cos(y)) / y ** 3


def mode_squares_histogram(D):
    """Number of grid modes with each integer |n|^2 on the D^3 FFT grid.

    Each axis holds n in {-(D/2-1), ..., D/2}; the three-axis histogram is
    built from the unique per-axis squares and their multiplicities.
    """
    D = int(D)
    h = D // 2
    sq = np.arange(h + 1, dtype=np.int64) ** 2
    mult = np.full(h + 1, 2, dtype=np.int64)
    mult[0] = 1
    mult[h] = 1
    n2max = 3 * h * h
    plane = np.bincount((sq[:, None] + sq[None, :]).ravel(), weights=(mult[:, None] * mult[None, :]).ravel(),
                        minlength=2 * h * h + 1)
    out = np.zeros(n2max + 1, dtype=np.float64)
    for s, m in zip(sq, mult):
        out[s:s + plane.size] += m * plane
    return np.rint(out).astype(np.int64)


class PowerLawSpectrum:
    """P(k) = amp k^n_index exp(-(k/k_cut)^2) for a box of side D cells and box_len Mpc."""

    def __init__(self, D, box_len, sigma_cell2=0.36, n_index=-2.5, k_cut=None):
        self.D = int(D)
        self.box_len = float(box_len)
        self.sigma_cell2 = float(sigma_cell2)
        self.n_index = float(n_index)
        self.k_cut = float(np.pi * self.D / self.box_len) if k_cut is None else float(k_cut)
        self.dk = 2.0 * np.pi / self.box_len
        self.volume = self.box_len ** 3
        hist = mode_squares_histogram(self.D)
        m = np.arange(1, hist.size)
        self._m = m
        self._hist = hist[1:]
        k = self.dk * np.sqrt(m.astype(np.float64))
        self._k = k
        self._shape_k = self.shape(k)
        self.amp = self.sigma_cell2 * self.volume / float(np.sum(self._hist * self._shape_k))
        self._s0 = {}  # S(R) by radius: the spectrum does not change after construction

    def shape(self, k):
        k = np.asarray(k, dtype=np.float64)
        out = np.zeros_like(k)
        pos = k > 0
        out[pos] = k[pos] ** self.n_index * np.exp(-(k[pos] / self.k_cut) ** 2)
        return out

    def pk(self, k):
        return self.amp * self.shape(k)

    def del2(self, k):
        k = np.asarray(k, dtype=np.float64)
        return k ** 3 * self.pk(k) / (2.0 * np.pi ** 2)

    def sig0_discrete(self, R):
        R = float(R)
        s = self._s0.get(R)
        if s is None:
            w = tophat_window(self._k * R)
            s = self._s0[R] = float(np.sum(self._hist * self.amp * self._shape_k * w * w) / self.volume)
        return s

    def sig0_trapz(self, R, nint=1000):
        return sig0_trapz(R, self.del2, nint)

    def sig0(self, R, method="discrete"):
        if method == "discrete":
            return self.sig0_discrete(R)
        if method == "trapz":
            return self.sig0_trapz(R)
        raise ValueError(f"unknown sig0 method {method!r}")


@functools.lru_cache(maxsize=32)
def power_law_spectrum(D, box_len, sigma_cell2=0.36, n_index=-2.5):
    """A shared PowerLawSpectrum per (D, box_len, sigma_cell2, n_index): its
    normalisation and its S(R) values are computed once per process."""
    return PowerLawSpectrum(int(D), float(box_len), float(sigma_cell2), float(n_index))


def sig0_trapz(R, del2_fn, nint=1000, factor=1.0e-6):
    """The reference's S(R) rule for a Delta^2(k) callable (sigmas.py:176-191)."""
    logk = np.arange(-20.0, 20.0, 0.1)
    k = np.exp(logk)
    vals = np.abs(del2_fn(k) * tophat_window_closed(float(R) * k) ** 2)
    keep = vals > vals.max() * factor
    lk = np.linspace(logk[keep].min(), logk[keep].max(), nint)
    kk = np.exp(lk)
    return float(np.trapezoid(del2_fn(kk) * tophat_window_closed(float(R) * kk) ** 2, lk))


def sig0(RL, spectrum, method="discrete"):
    """S(R) for a spectrum object (scalar or array R), as sigmas.py:224-232 returns."""
    if np.isscalar(RL):
        return spectrum.sig0(float(RL), method)
    return np.array([spectrum.sig0(float(r), method) for r in np.ravel(RL)]).reshape(np.shape(RL))
