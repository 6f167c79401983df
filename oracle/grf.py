"""Gaussian random field initial conditions (initialize stage) on the CPU.

TEST INFRASTRUCTURE ONLY (see ``oracle/__init__.py``).

The reference (Programs/initialize.py:33-111, initialize.cu:58-95,133-190)
draws delta_k = sqrt(V P(k)) (g1 + i g2) independently at every k and keeps the
real part of the inverse transform.  That is distributed exactly like a
Hermitian spectrum with E|delta_k|^2 = V P(k): paired modes with real and
imaginary parts ~ N(0, VP/2), self-conjugate modes real ~ N(0, VP), DC = 0.
This module builds that Hermitian half spectrum directly, from Philox4x32-10
keyed by (seed, box id) and counted by the linear index of each mode's
canonical representative (see ``grf_half_spectrum``).

Conventions: the stored coefficients are ``c2r-ready``: an unnormalised
inverse real transform of them gives delta(x) = (1/V) sum_k delta_k e^{ikx},
i.e. the reference's Re[IFFT_(1/N)(delta_k)] / scale^3 (initialize.py:80-81).
"""
import numpy as np

from .philox import philox4x32_10, box_muller_f64
from .sigma import mode_index, synthetic_amplitude, synthetic_pk_shape
from .bubbles import kmag_half_f32, window_f32, apply_window_f32, delta_k_f32


def table_pk(k, k_table, p_table):
    """Linear interpolation of a float32 (K, P) table, initialize.cu:87-90
    (i = first index with K[i] >= k; 0 outside the table), in FP64."""
    K = np.asarray(k_table, dtype=np.float32).astype(np.float64)
    Pt = np.asarray(p_table, dtype=np.float32).astype(np.float64)
    return np.interp(np.asarray(k, dtype=np.float64), K, Pt, left=0.0, right=0.0)


def grf_half_spectrum(D, box_len, seed, box_id=0, sigma_cell2=0.36, pk_table=None):
    """complex64 (D, D, D//2+1) c2r-ready Hermitian GRF coefficients (FP64 math).

    pk_table: None for the synthetic power law of the benchmark recipe, or a
    float32 (K, P) pair interpolated as the reference's init_kernel does."""
    H = D // 2 + 1
    a = np.arange(D)[:, None, None]
    b = np.arange(D)[None, :, None]
    c = np.arange(H)[None, None, :]
    a, b, c = np.broadcast_arrays(a, b, c)
    ma, mb = (D - a) % D, (D - b) % D
    edge = (c == 0) | (c == D // 2)
    lin_self = a * D + b
    lin_mirror = ma * D + mb
    use_mirror = edge & (lin_mirror < lin_self)
    ra = np.where(use_mirror, ma, a)
    rb = np.where(use_mirror, mb, b)
    rep = ((ra.astype(np.int64) * D + rb) * H + c).ravel()
    ctr = np.zeros((rep.size, 4), dtype=np.uint32)
    ctr[:, 0] = (rep & 0xFFFFFFFF).astype(np.uint32)
    ctr[:, 1] = (rep >> 32).astype(np.uint32)
    key = np.array([seed & 0xFFFFFFFF, box_id & 0xFFFFFFFF], dtype=np.uint32)
    u = philox4x32_10(ctr, key)
    g1, g2 = box_muller_f64(u[:, 0], u[:, 1])
    g1 = g1.reshape(a.shape)
    g2 = g2.reshape(a.shape)

    dk = 2.0 * np.pi / float(box_len)
    n = mode_index(D).astype(np.float64)
    m2 = n[a] ** 2 + n[b] ** 2 + c.astype(np.float64) ** 2
    k = dk * np.sqrt(m2)
    V = float(box_len) ** 3
    if pk_table is None:
        A = synthetic_amplitude(D, box_len, sigma_cell2)
        amp = np.sqrt(A * synthetic_pk_shape(k, np.pi * D / float(box_len)) / V)
    else:
        amp = np.sqrt(table_pk(k, *pk_table) / V)

    selfconj = edge & (ma == a) & (mb == b)
    re = np.where(selfconj, amp * g1, amp * g1 / np.sqrt(2.0))
    im = np.where(selfconj, 0.0, amp * g2 / np.sqrt(2.0))
    im = np.where(use_mirror, -im, im)
    out = (re + 1j * im).astype(np.complex64)
    out[0, 0, 0] = 0
    return out


def c2r(spec, D):
    """Unnormalised inverse real transform: sum_k spec_k e^{ikx} (FP64)."""
    return np.fft.irfftn(spec.astype(np.complex128), s=(D, D, D), axes=(0, 1, 2)) * float(D) ** 3


def smoothing_radius(box_len, hii_dim):
    """smoothR = float32(L_FACTOR * BOX_LEN / HII_DIM), initialize.py:87."""
    from .sigma import L_FACTOR
    return np.float32(float(np.float32(L_FACTOR) * np.float32(box_len)) / float(hii_dim))


def subsample_indices(sw, pixel_factor):
    """Nearest pick l = floor(k*pixel_factor + 0.5), initialize.cu:140-142
    (k*pixel_factor in float32, + 0.5 in double)."""
    k = np.arange(sw, dtype=np.float32)
    return np.floor((k * np.float32(pixel_factor)).astype(np.float32).astype(np.float64) + 0.5).astype(np.int64)


def subsample(large, sw):
    """initialize.cu:133-146: small[i,j,k] = large[l(i), l(j), l(k)]."""
    pf = np.float32(large.shape[0]) / np.float32(sw)
    l = subsample_indices(sw, pf)
    return np.ascontiguousarray(large[np.ix_(l, l, l)])


def velocity_multiplier(D, box_len, comp):
    """i k_c / k^2 on the half grid (initialize.cu:156-189), FP64; zero at k = 0
    and on the Nyquist plane of axis ``comp`` (whose correlated part the
    reference's real-part projection cancels)."""
    dk = 2.0 * np.pi / float(box_len)
    n = mode_index(D).astype(np.float64)
    nh = np.arange(D // 2 + 1, dtype=np.float64)
    n0 = n[:, None, None]
    n1 = n[None, :, None]
    n2 = nh[None, None, :]
    ksq = dk * dk * (n0 ** 2 + n1 ** 2 + n2 ** 2)
    kc = dk * [n0, n1, n2][comp]
    idx = [np.arange(D)[:, None, None], np.arange(D)[None, :, None],
           np.arange(D // 2 + 1)[None, None, :]][comp]
    with np.errstate(divide="ignore", invalid="ignore"):
        f = np.where(ksq > 0, kc / ksq, 0.0)
    f = np.where(idx == D // 2, 0.0, f)
    return 1j * np.broadcast_to(f, (D, D, D // 2 + 1))


def initialize(D, hii_dim, box_len, seed, box_id=0, sigma_cell2=0.36, with_velocity=True, pk_table=None):
    """Initial conditions as initialize.init (initialize.py:33-111).

    Returns dict with ``deltax`` (D^3), ``smoothed`` (hii_dim^3) and, when asked,
    ``v`` = (vx, vy, vz) at hii_dim^3, all float32.  Smoothing uses the type-0
    fall-through window at smoothR with the slowest-axis-first |k| order of
    initialize.cu:24; the windows and the final casts are float32, the
    transforms FP64.
    """
    spec = grf_half_spectrum(D, box_len, seed, box_id, sigma_cell2, pk_table)
    out = {"spec": spec, "deltax": c2r(spec, D).astype(np.float32)}
    R = smoothing_radius(box_len, hii_dim)
    kmag = kmag_half_f32(D, delta_k_f32(box_len), fastest_first=False)
    w0, cut, w2 = window_f32(kmag, R, 0, True)
    sm = apply_window_f32(spec, w0, cut, w2)
    out["smoothed"] = subsample(c2r(sm, D).astype(np.float32), hii_dim)
    if with_velocity:
        vs = []
        for comp in range(3):
            v = (spec.astype(np.complex128) * velocity_multiplier(D, box_len, comp)).astype(np.complex64)
            v = apply_window_f32(v, w0, cut, w2)
            vs.append(subsample(c2r(v, D).astype(np.float32), hii_dim))
        out["v"] = tuple(vs)
    return out
