"""Pins the CPU oracle against known answers, the reference's own data tables,
closed forms, library routines and brute-force loop restatements."""
import os

import numpy as np
import pytest
from scipy.special import erfc, spherical_jn

from conftest import GOLDEN
from oracle import bubbles, deltat, grf, perturb, philox, sigma


# ----------------------------------------------------------------- Philox

def test_philox_random123_kat():
    # Known-answer vectors of Philox4x32-10 published with Random123.
    cases = [
        ((0, 0, 0, 0), (0, 0), (0x6627E8D5, 0xE169C58D, 0xBC57AC4C, 0x9B00DBD8)),
        ((0xFFFFFFFF,) * 4, (0xFFFFFFFF,) * 2, (0x408F276D, 0x41C83B0E, 0xA20BC7C6, 0x6D5451FD)),
        ((0x243F6A88, 0x85A308D3, 0x13198A2E, 0x03707344), (0xA4093822, 0x299F31D0),
         (0xD16CFE09, 0x94FDCCEB, 0x5001E420, 0x24126EA1)),
    ]
    for ctr, key, want in cases:
        got = philox.philox4x32_10(np.array(ctr, np.uint32), np.array(key, np.uint32))
        assert tuple(int(x) for x in got) == want


def test_philox_matches_rocrand_engine():
    # tests/golden/philox_rocrand.txt: rocRAND's philox4x32_10_engine compiled
    # for the host; words c*4..c*4+3 are Philox(counter=(c,0,sub_lo,sub_hi), key=seed).
    rows = open(os.path.join(GOLDEN, "philox_rocrand.txt")).read().split("\n")
    n = 0
    for line in rows:
        if not line.strip():
            continue
        v = [int(x) for x in line.split()]
        seed, sub, words = v[0], v[1], v[2:]
        key = np.array([seed & 0xFFFFFFFF, seed >> 32], np.uint32)
        for c in range(2):
            ctr = np.array([c, 0, sub & 0xFFFFFFFF, sub >> 32], np.uint32)
            assert [int(x) for x in philox.philox4x32_10(ctr, key)] == words[4 * c:4 * c + 4]
        n += 1
    assert n == 12


def test_box_muller_distribution():
    from scipy import stats
    u = philox.philox4x32_10(np.stack([np.arange(200000, dtype=np.uint32)] + [np.zeros(200000, np.uint32)] * 3, 1),
                             np.array([5, 9], np.uint32))
    a = philox.uniform_open01(u[:, 0])
    assert a.min() > 0 and a.max() < 1
    g1, g2 = philox.box_muller_f64(u[:, 0], u[:, 1])
    g = np.concatenate([g1, g2])
    assert abs(g.mean()) < 5 / np.sqrt(g.size)
    assert abs(g.var() - 1) < 5 * np.sqrt(2 / g.size)
    assert stats.kstest(g, "norm").pvalue > 1e-3
    assert abs(np.corrcoef(g1, g2)[0, 1]) < 5 / np.sqrt(g1.size)


# ----------------------------------------------------------------- schedule and S(R)

@pytest.mark.parametrize("L,D,count", [(100, 64, 37), (300, 256, 40), (500, 512, 41), (150, 96, 37),
                                       (200, 128, 37), (250, 160, 37)])
def test_radius_schedule_counts(L, D, count):
    r, cell = sigma.radius_schedule(L, D)
    assert len(r) == count and r.dtype == np.float32
    assert r[0] == np.float32(30.0) and r[1] == np.float32(27.272728) and r[2] == np.float32(24.793388)
    stop = float(sigma.L_FACTOR) * cell
    # the last radius is the first whose successor falls below the stop rule
    assert float(r[-1]) / 1.1 <= max(stop, 0.5)
    assert float(r[-2]) / 1.1 > max(stop, 0.5)
    assert np.all(np.diff(r) < 0)


def test_radius_schedule_py2_integer_cell():
    r, cell = sigma.radius_schedule(300, 256, py2_int_scale=True)
    assert cell == 1
    assert float(r[-1]) / 1.1 <= 0.620350491 and float(r[-2]) / 1.1 > 0.620350491


def test_sig0_trapz_matches_reference_table():
    t = np.load(os.path.join(GOLDEN, "reference_tables.npz"))
    f = sigma.table_del2(t["k"], t["Del2"])
    for R, S in zip(t["radius"], t["sig0"]):
        assert abs(sigma.sig0_trapz(float(R), f) / S - 1) < 2e-4, R
    # sigma_8 of the Planck13 table: S(8/h Mpc) = 0.829^2 within the trapz error
    assert abs(sigma.sig0_trapz(8 / 0.673, f) / 0.829 ** 2 - 1) < 5e-3


def test_sig0_discrete_bruteforce_and_limits():
    D, L = 12, 30.0
    A = sigma.synthetic_amplitude(D, L)
    n = sigma.mode_index(D)
    nx, ny, nz = np.meshgrid(n, n, n, indexing="ij")
    k = 2 * np.pi / L * np.sqrt(nx ** 2 + ny ** 2 + nz ** 2).ravel()
    k = k[k > 0]
    P = A * k ** -2.5 * np.exp(-(k / (np.pi * D / L)) ** 2)
    assert abs(P.sum() / L ** 3 - 0.36) < 1e-12
    for R in (0.7, 2.0, 5.5):
        x = k * R
        W = 3 * spherical_jn(1, x) / x
        assert abs(np.sum(P * W ** 2) / L ** 3 - sigma.sig0_discrete(R, D, L)) < 1e-12
    assert abs(sigma.sig0_discrete(1e-7, D, L) - 0.36) < 1e-9
    vals = [sigma.sig0_discrete(R, D, L) for R in (0.5, 1, 2, 4, 8)]
    assert np.all(np.diff(vals) < 0)


def test_nsq_histogram_counts_every_mode():
    for D in (4, 10, 16):
        h = sigma.nsq_histogram_bruteforce(D)
        assert h.sum() == D ** 3
        n = sigma.mode_index(D)
        assert h[1] == 6 and h[0] == 1
        assert set(n) == set(range(-(D // 2) + 1, D // 2 + 1))


# ----------------------------------------------------------------- filter windows

def test_windows_closed_forms():
    k = np.linspace(0, 3.0, 2001).astype(np.float32)
    R = np.float32(1.7)
    x = k.astype(np.float64) * float(R)
    # type 2, textbook: Gaussian exp(-(0.643 kR)^2/2)
    w0, cut, w2 = bubbles.window_f32(k, R, 2, ref_fallthrough=False)
    assert not cut.any() and np.all(w0 == 1)
    assert np.allclose(w2, np.exp(-(0.643 * x) ** 2 / 2), rtol=2e-6, atol=1e-7)
    # type 0, textbook: W_TH = 3 j1(x)/x
    # (the reference's float32 form sin/x^3 - cos/x^2 cancels for small x:
    #  pinned to the closed form for x > 0.3, bounded below)
    w0, cut, w2 = bubbles.window_f32(k[1:], R, 0, ref_fallthrough=False)
    wth = 3 * spherical_jn(1, x[1:]) / x[1:]
    big = x[1:] > 0.3
    assert np.allclose(w0[big], wth[big], rtol=1e-4, atol=2e-6)
    assert np.all(np.abs(w0[~big] - wth[~big]) < 0.05)
    # reference chain for type 1: cut at kR > 1/0.413566994, Gaussian of 0.413566994*0.643 kR
    w0, cut, w2 = bubbles.window_f32(k, R, 1, ref_fallthrough=True)
    assert np.array_equal(cut, x * 0.413566994 > 1 + 1e-6) or np.sum(cut != (x * 0.413566994 > 1)) <= 1
    assert np.allclose(w2, np.exp(-(0.413566994 * 0.643 * x) ** 2 / 2), rtol=2e-6, atol=1e-7)
    # reference chain for type 0 multiplies the top-hat by the type-1 chain
    w0, cut0, w20 = bubbles.window_f32(k[1:], R, 0, ref_fallthrough=True)
    assert np.array_equal(cut0, cut[1:]) and np.array_equal(w20, w2[1:])
    # f64 flavour agrees with the product of the f32 pieces
    w64 = bubbles.window_f64(k[1:].astype(np.float64), float(R), 0, True)
    assert np.allclose(np.where(cut0, 0, w0.astype(np.float64) * w20)[big], w64[big], rtol=1e-5, atol=1e-6)


def _full_kmag_f32(D, dk):
    n = sigma.mode_index(D).astype(np.float32)
    kk = (n * np.float32(dk)).astype(np.float32)
    # find_bubbles.cu INDEX layout: i (axis 2) is k_x; sum (kx^2 + ky^2) + kz^2
    kx = kk[None, None, :]
    ky = kk[None, :, None]
    kz = kk[:, None, None]
    s = ((kx * kx).astype(np.float32) + (ky * ky).astype(np.float32)).astype(np.float32)
    s = (s + (kz * kz).astype(np.float32)).astype(np.float32)
    return np.sqrt(s).astype(np.float32)


@pytest.mark.parametrize("fil", [0, 1, 2])
def test_filter_half_spectrum_equals_reference_c2c(fil):
    # The reference filters the full c2c spectrum and keeps the real part of
    # the inverse (find_HII_bubbles.py:141-156).  The oracle filters the r2c
    # half spectrum.  Compare on a D=8 box with a brute-force DFT.
    D, L, R = 8, 20.0, np.float32(3.1)
    rng = np.random.default_rng(1)
    x = rng.standard_normal((D, D, D)).astype(np.float32)
    dk = bubbles.delta_k_f32(L)
    idx = np.arange(D)
    E = np.exp(-2j * np.pi * np.outer(idx, idx) / D)
    F = np.einsum("ai,bj,ck,ijk->abc", E, E, E, x.astype(np.complex128))
    w0, cut, w2 = bubbles.window_f32(_full_kmag_f32(D, dk), R, fil, True)
    Ff = np.where(cut, 0, F * w0 * w2)
    Einv = np.conj(E)
    back = np.einsum("ai,bj,ck,ijk->abc", Einv, Einv, Einv, Ff).real / D ** 3
    half = np.fft.rfftn(x.astype(np.float64))
    w0h, cuth, w2h = bubbles.window_f32(bubbles.kmag_half_f32(D, dk), R, fil, True)
    mine = np.fft.irfftn(np.where(cuth, 0, half * w0h * w2h), s=(D, D, D), axes=(0, 1, 2))
    assert np.allclose(mine, back, atol=1e-6 * np.abs(back).max() + 1e-12)


# ----------------------------------------------------------------- find_bubbles loop

def _loop_python(delta, L, radii, denoms, zeta, deltac):
    """Pure-Python restatement of the centre-pixel loop with the reference's
    full c2c transforms and per-cell kernels (find_bubbles.cu:43-71,142-158)."""
    D = delta.shape[0]
    dk = bubbles.delta_k_f32(L)
    km = _full_kmag_f32(D, dk)
    F = np.fft.fftn(delta.astype(np.complex128))
    ion = np.zeros_like(delta)
    for r, (R, den) in enumerate(zip(radii, denoms)):
        w0, cut, w2 = bubbles.window_f32(km, R, 1, True)
        sm = np.fft.ifftn(np.where(cut, 0, F * w0 * w2)).real.astype(np.float32)
        final = r == len(radii) - 1
        for p in np.ndindex(D, D, D):
            if ion[p] == 1:
                continue
            num = np.float32(np.float32(deltac) - sm[p])
            fc = np.float32(erfc(float(np.float32(num / np.float32(den)))))
            if float(fc) >= 1.0 / zeta:
                ion[p] = 1.0
            elif final:
                ion[p] = np.float32(float(fc) * zeta)
    return ion


def test_loop_matches_python_restatement_tiny():
    D, L = 8, 12.5
    delta = grf.c2r(grf.grf_half_spectrum(D, L, seed=3), D).astype(np.float32)
    radii, _ = sigma.radius_schedule(L, D)
    den = sigma.denominators(radii, 0.5, lambda R: sigma.sig0_discrete(R, D, L))
    want = _loop_python(delta, L, radii, den, 30.0, 1.686)
    got = bubbles.find_bubbles_f32(delta, L, radii, den, 30.0, 1.686)
    assert np.array_equal(got == 1, want == 1)
    assert np.allclose(got, want, rtol=1e-5, atol=1e-7)
    assert 0 < (got == 1).mean() < 1


def test_constant_box_closed_form():
    D, L = 16, 30.0
    radii, _ = sigma.radius_schedule(L, D)
    den = sigma.denominators(radii, 0.6, lambda R: sigma.sig0_discrete(R, D, L))
    for c, zeta in ((0.2, 3.0), (-0.9, 2.0)):
        delta = np.full((D, D, D), c, np.float32)
        got = bubbles.find_bubbles_f32(delta, L, radii, den, zeta, 1.686)
        f = erfc((1.686 - c) / den.astype(np.float64))
        if np.any(zeta * f >= 1):
            assert np.all(got == 1)
        else:
            assert np.allclose(got, zeta * f[-1], rtol=1e-6)


def test_monotone_in_zeta_and_no_reversion():
    d = np.load(os.path.join(GOLDEN, "bubbles_D16.npz"))
    args = (d["delta"], float(d["box_len"]), d["radii"], d["denoms"])
    lo = bubbles.find_bubbles_f32(*args, 20.0, float(d["deltac"]))
    hi = bubbles.find_bubbles_f32(*args, 40.0, float(d["deltac"]))
    assert np.all(hi[lo == 1] == 1)
    assert (hi == 1).sum() > (lo == 1).sum()
    _, trace = bubbles.find_bubbles_f32(*args, 30.0, float(d["deltac"]), return_trace=True)
    hit_any = np.zeros(d["delta"].shape, bool)
    for t in trace[:-1]:
        hit_any |= t["fcoll"].astype(np.float64) >= 1 / 30.0
    ion = bubbles.find_bubbles_f32(*args, 30.0, float(d["deltac"]))
    assert np.array_equal(ion == 1, hit_any | (trace[-1]["fcoll"].astype(np.float64) >= 1 / 30.0))


@pytest.mark.parametrize("name", ["bubbles_D16.npz", "bubbles_D32_nl.npz"])
def test_loop_golden_vectors(name):
    d = np.load(os.path.join(GOLDEN, name))
    LE = bool(d["LE"])
    got = bubbles.find_bubbles_f32(d["delta"], float(d["box_len"]), d["radii"], d["denoms"], float(d["zeta"]),
                                   float(d["deltac"]), LE=LE, growth_inv=np.float32(float(d["deltac"]) / 1.686),
                                   fc_mean=d["fc_mean"])
    assert np.array_equal(got, d["ionized"])
    ref64, margin = bubbles.find_bubbles_f64(d["delta"], float(d["box_len"]), d["radii"], d["denoms"],
                                             float(d["zeta"]), float(d["deltac"]), LE=LE,
                                             growth_inv=float(d["deltac"]) / 1.686, fc_mean=float(d["fc_mean"]))
    ok = margin > 1e-5
    assert np.array_equal((got == 1)[ok], (ref64 == 1)[ok])


def _paint_bruteforce(centers, R, periodic):
    D = centers.shape[0]
    out = np.zeros(centers.shape, np.float32)
    R = np.float32(R)
    R2 = np.float32(R * R)
    for (k, j, i) in zip(*np.nonzero(centers)):
        if periodic:
            rng = [range(int(np.floor(c - R)), int(np.ceil(c + R))) for c in (k, j, i)]
        else:  # find_bubbles.cu:87-92, clipped bounding box
            rng = [range(0 if c - R < 0 else int(np.floor(c - R)), D if c + R > D else int(np.ceil(c + R)))
                   for c in (k, j, i)]
        for kk in rng[0]:
            for jj in rng[1]:
                for ii in rng[2]:
                    rsq = np.float32((ii - i) ** 2 + (jj - j) ** 2 + (kk - k) ** 2)
                    if rsq < R2:
                        out[kk % D, jj % D, ii % D] = 1.0
    return out


@pytest.mark.parametrize("periodic", [False, True])
@pytest.mark.parametrize("R", [1.0, 2.5, 3.3, 7.9])
def test_sphere_paint_bruteforce(periodic, R):
    rng = np.random.default_rng(int(R * 10) + periodic)
    centers = rng.random((12, 12, 12)) < 0.01
    centers[0, 0, 0] = centers[11, 5, 6] = True
    got = bubbles.paint_spheres(np.zeros((12, 12, 12), np.float32), centers, np.float32(R), periodic)
    assert np.array_equal(got, _paint_bruteforce(centers, R, periodic))


# ----------------------------------------------------------------- initial conditions

def test_grf_hermitian_and_variance():
    D, L = 16, 40.0
    s = grf.grf_half_spectrum(D, L, seed=11, box_id=2)
    x = grf.c2r(s, D)
    assert np.allclose(np.fft.rfftn(x) / D ** 3, s, atol=1e-7 * np.abs(s).max())
    assert s[0, 0, 0] == 0
    # self-conjugate modes are real
    for a in (0, D // 2):
        for b in (0, D // 2):
            for c in (0, D // 2):
                assert s[a, b, c].imag == 0
    # E|delta_k|^2 = P/V per mode: average over the modes of the box
    n = sigma.mode_index(D)
    m2 = n[:, None, None] ** 2 + n[None, :, None] ** 2 + np.arange(D // 2 + 1)[None, None, :] ** 2
    P = sigma.synthetic_pk(2 * np.pi / L * np.sqrt(m2.astype(float)), D, L)
    ratio = (np.abs(s.astype(np.complex128)) ** 2)[m2 > 0] / (P[m2 > 0] / L ** 3)
    assert abs(ratio.mean() - 1) < 0.05
    many = [grf.c2r(grf.grf_half_spectrum(D, L, seed=q), D).var() for q in range(20)]
    assert abs(np.mean(many) - 0.36) < 0.04
    assert not np.array_equal(s, grf.grf_half_spectrum(D, L, seed=12, box_id=2))
    assert np.array_equal(s, grf.grf_half_spectrum(D, L, seed=11, box_id=2))


def test_subsample_indices():
    assert np.array_equal(grf.subsample_indices(64, 2.0), 2 * np.arange(64))
    pf = np.float32(512) / np.float32(96)
    want = [int(np.floor(float(np.float32(np.float32(k) * pf)) + 0.5)) for k in range(96)]
    assert list(grf.subsample_indices(96, pf)) == want
    big = np.arange(16 ** 3, dtype=np.float32).reshape(16, 16, 16)
    assert np.array_equal(grf.subsample(big, 8), big[::2, ::2, ::2])


def test_velocity_is_gradient_of_inverse_laplacian():
    # psi_c = d/dx_c of phi with laplacian(phi) = -delta, so div(psi) = -delta
    # for the modes the Nyquist planes do not touch.
    D, L = 16, 32.0
    s = grf.grf_half_spectrum(D, L, seed=4)
    n = sigma.mode_index(D)
    keep = (np.abs(n) < D // 2)
    s = s * keep[:, None, None] * keep[None, :, None] * (np.arange(D // 2 + 1) < D // 2)[None, None, :]
    delta = grf.c2r(s, D)
    div = 0
    for comp in range(3):
        psi_k = s * grf.velocity_multiplier(D, L, comp)
        kc = 2 * np.pi / L * [n[:, None, None], n[None, :, None], np.arange(D // 2 + 1)[None, None, :]][comp]
        div = div + grf.c2r(1j * kc * psi_k, D)
    assert np.allclose(div, -delta, atol=1e-9 * np.abs(delta).max())


# ----------------------------------------------------------------- perturb_field

def _ngp_python(DIM, vx, vy, vz, s):
    """Pure-Python perturb_field.cu:100-127 per particle."""
    sw = vx.shape[0]
    swf = np.float32(sw)
    pf = float(np.float32(DIM) / np.float32(sw))
    out = {}
    for i in range(DIM):
        for j in range(DIM):
            for k in range(DIM):
                hi, hj, hk = int(np.floor(i / pf)), int(np.floor(j / pf)), int(np.floor(k / pf))
                pos = []
                for c, v in ((i, vx), (j, vy), (k, vz)):
                    f = np.float32((c + 0.5) / DIM)
                    f = np.float32(f + np.float32(v[hi, hj, hk] * np.float32(s)))
                    f = np.float32(f * swf)
                    while f >= swf:
                        f = np.float32(f - swf)
                    while f < 0:
                        f = np.float32(f + swf)
                    pos.append(int(f))
                out[(i, j, k)] = tuple(pos)
    return out


def test_ngp_targets_match_python_loop():
    rng = np.random.default_rng(2)
    DIM, sw = 8, 4
    v = [rng.standard_normal((sw, sw, sw)).astype(np.float32) * 30 for _ in range(3)]
    s = perturb.displacement_scale(0.12, 0.0036, 25.0)
    xi, yi, zi = perturb.ngp_targets_f32(DIM, *v, s)
    want = _ngp_python(DIM, *v, s)
    for key, (a, b, c) in want.items():
        assert (xi[key], yi[key], zi[key]) == (a, b, c)


def test_move_mass_invariants():
    rng = np.random.default_rng(3)
    DIM, sw, L = 16, 8, 40.0
    delta = rng.standard_normal((DIM,) * 3).astype(np.float32)
    zero = np.zeros((sw,) * 3, np.float32)
    # no displacement: NGP block sums of (1 + D_init delta), renormalised
    got = perturb.move_mass_ngp(delta, zero, zero, zero, 0.1, 0.1, L)
    m = perturb.particle_mass_f32(delta, 0.1).astype(np.float64)
    blocks = m.reshape(sw, 2, sw, 2, sw, 2).sum(axis=(1, 3, 5))
    assert np.allclose(got, blocks / 8 - 1, atol=1e-6)
    # a shift of exactly one HII cell along axis 0 rolls the result
    one = np.full((sw,) * 3, np.float32(1.0 / sw), np.float32)
    shifted = perturb.move_mass_ngp(delta, one, zero, zero, 1.0, 0.0, 1.0)
    base = perturb.move_mass_ngp(delta, zero, zero, zero, 1.0, 0.0, 1.0)
    assert np.allclose(shifted, np.roll(base, 1, axis=0), atol=1e-6)
    # mass conservation for NGP and CIC with real displacements
    v = [rng.standard_normal((sw,) * 3).astype(np.float32) * 20 for _ in range(3)]
    for fn in (perturb.move_mass_ngp, perturb.move_mass_cic):
        out = fn(delta, *v, 0.12, 0.0036, L).astype(np.float64)
        tot = perturb.particle_mass_f32(delta, 0.0036).astype(np.float64).sum()
        assert abs((out + 1).sum() * 8 / tot - 1) < 1e-6
    # CIC at pf = 1 with no displacement puts each particle on its own cell centre
    d8 = delta[:8, :8, :8].copy()
    z8 = np.zeros((8,) * 3, np.float32)
    assert np.allclose(perturb.move_mass_cic(d8, z8, z8, z8, 0.5, 0.5, L),
                       perturb.move_mass_ngp(d8, z8, z8, z8, 0.5, 0.5, L), atol=1e-6)


# ----------------------------------------------------------------- delta_T and P(k)

def test_delta_T_basic():
    xH = np.ones((4, 4, 4), np.float32)
    d = np.zeros((4, 4, 4), np.float32)
    c = deltat.tb_constant(8.0)
    assert np.all(deltat.delta_T_f32(xH, d, 8.0) == c)
    assert np.all(deltat.delta_T_f32(0 * xH, d + 3, 8.0) == 0)
    assert 20 < c < 30  # ~ 25 mK at z = 8 for Planck13 baryons


def test_power_spectrum_plane_wave_and_counts():
    D, L = 32, 64.0
    edges = deltat.log_edges(L, D, 12)
    x = np.arange(D)
    n0 = 3
    A = 0.2
    field = 1 + A * np.cos(2 * np.pi * n0 * x / D)[None, None, :] * np.ones((D, D, 1))
    k_ave, ps_ave, cnt, _ = deltat.power_spectrum(field, L, edges)
    k0 = 2 * np.pi / L * n0
    b = np.searchsorted(edges, k0)
    V = L ** 3
    assert abs(ps_ave[b] * cnt[b] / 2 - k0 ** 3 * (A * V / 2) ** 2 / (2 * np.pi ** 2 * V)) < 1e-9 * ps_ave[b] * cnt[b]
    others = np.delete(ps_ave * cnt, b)
    assert np.all(np.abs(others) < 1e-12 * ps_ave[b] * cnt[b])
    # counts: brute force over the full grid
    n = sigma.mode_index(D)
    kk = 2 * np.pi / L * np.sqrt(n[:, None, None] ** 2 + n[None, :, None] ** 2 + n[None, None, :] ** 2).ravel()
    kk = kk[kk > 0]
    bb = np.searchsorted(edges, kk, side="left")
    want = np.bincount(bb[bb < len(edges)], minlength=len(edges))
    assert np.array_equal(cnt, want)
