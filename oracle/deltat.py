"""21-cm brightness temperature and its spherically binned power spectrum.

TEST INFRASTRUCTURE ONLY (see ``oracle/__init__.py``).

``delta_T_f32`` restates Programs/delta_T.py:73-75.  ``power_spectrum`` keeps
the conventions of delta_T.py:80-98 and delta_T.cu:5-29 (unnormalised forward
transform of (dT/<dT> - 1) V/N, Delta^2 = k^3 |d_k|^2 / (2 pi^2 V), bin index =
first edge K[c] >= |k|) with the corrections: true box dimension, DC excluded,
modes above the last edge dropped, FP64 sums.  Half-grid modes with
0 < k_z < Nyquist stand for two modes (k and -k).
"""
import numpy as np

from .sigma import mode_index

PLANCK13 = dict(omega_M_0=0.315, omega_b_0=0.0487, h=0.673)  # INIT_PARAMS.py:33-42


def tb_constant(z, omega_b_0=0.0487, h=0.673, omega_M_0=0.315):
    """float32(27 (Ob h^2/0.023) sqrt((0.15/Om/h^2)(1+z)/10)), delta_T.py:73-74."""
    return np.float32(27 * (omega_b_0 * h * h / 0.023) *
                      np.sqrt((0.15 / omega_M_0 / h / h) * (1 + z) / 10.0))


def delta_T_f32(xH, delta, z, **cosmo):
    c = tb_constant(z, **cosmo)
    return np.asarray(np.asarray(c * np.asarray(xH, np.float32), np.float32) *
                      np.asarray(1.0 + np.asarray(delta, np.float32), np.float32), np.float32)


def log_edges(box_len, D, nbins=40):
    """nbins log-spaced upper edges from k_fund = 2 pi/L to k_nyq = pi D / L."""
    return np.logspace(np.log10(2 * np.pi / box_len), np.log10(np.pi * D / box_len), nbins)


def bin_of_nsq(nsq, box_len, edges):
    """Bin index of an integer |n|^2: first c with edges[c] >= dk sqrt(nsq);
    len(edges) when above the last edge; -1 for the DC mode."""
    dk = 2.0 * np.pi / float(box_len)
    k = dk * np.sqrt(np.asarray(nsq, dtype=np.float64))
    b = np.searchsorted(np.asarray(edges, dtype=np.float64), k, side="left")
    return np.where(np.asarray(nsq) == 0, -1, b)


def power_spectrum(field, box_len, edges, dimensional=False):
    """Returns (k_ave, ps_ave, counts, sums) with sums = (ps_sum, k_sum) in FP64."""
    f = np.asarray(field, dtype=np.float64)
    D = f.shape[0]
    N = D ** 3
    V = float(box_len) ** 3
    ave = f.mean()
    dd = (f / ave - 1.0) * V / N
    if dimensional:
        dd = dd * ave
    F = np.fft.rfftn(dd)
    n = mode_index(D)
    nh = np.arange(D // 2 + 1)
    nsq = n[:, None, None] ** 2 + n[None, :, None] ** 2 + nh[None, None, :] ** 2
    w = np.where((nh == 0) | (nh == D // 2), 1, 2)[None, None, :] * np.ones_like(nsq)
    b = bin_of_nsq(nsq, box_len, edges)
    nb = len(edges)
    ok = (b >= 0) & (b < nb)
    dk = 2.0 * np.pi / float(box_len)
    k = dk * np.sqrt(nsq.astype(np.float64))
    contrib = k ** 3 * np.abs(F) ** 2 / (2 * np.pi ** 2 * V)
    cnt = np.bincount(b[ok], weights=w[ok], minlength=nb).astype(np.int64)
    ps = np.bincount(b[ok], weights=(w * contrib)[ok], minlength=nb)
    ks = np.bincount(b[ok], weights=(w * k)[ok], minlength=nb)
    with np.errstate(invalid="ignore", divide="ignore"):
        k_ave = np.where(cnt > 0, ks / np.maximum(cnt, 1), 0.0)
        ps_ave = np.where(cnt > 0, ps / np.maximum(cnt, 1), 0.0)
    return k_ave, ps_ave, cnt, (ps, ks)
