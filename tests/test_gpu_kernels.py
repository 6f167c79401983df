"""GPU parity of the individual cm21 kernels against the CPU oracle (through
the C ABI).  Tolerances: integer/index work and cut-off decisions bit-exact;
float32 fields within 1e-5 relative to the field's maximum unless stated."""
import numpy as np
import pytest
import torch

from conftest import pkg
from oracle import bubbles, deltat, grf, perturb, sigma

pytestmark = pytest.mark.gpu


def _rand_spec(D, seed):
    rng = np.random.default_rng(seed)
    s = (rng.standard_normal((D, D, D // 2 + 1)) + 1j * rng.standard_normal((D, D, D // 2 + 1))).astype(np.complex64)
    return s


def _dev(a):
    return torch.from_numpy(np.ascontiguousarray(a)).cuda()


@pytest.mark.parametrize("fil,fall,korder", [(0, True, 0), (1, True, 0), (2, True, 0), (0, False, 1), (1, False, 0),
                                             (2, False, 0), (0, True, 1)])
def test_filter_matches_oracle(gpu, fil, fall, korder):
    g = pkg("gpu")
    D, L, R = 32, 50.0, np.float32(4.7)
    s = _rand_spec(D, fil + 10 * korder)
    out = g.hii_filter(_dev(s), L, R, fil, fall, korder).cpu().numpy()
    km = bubbles.kmag_half_f32(D, bubbles.delta_k_f32(L), fastest_first=(korder == 0))
    w0, cut, w2 = bubbles.window_f32(km, R, fil, fall)
    want = bubbles.apply_window_f32(s, w0, cut, w2)
    assert np.array_equal(out == 0, want == 0), "cut-off decisions differ"
    assert np.allclose(out, want, rtol=2e-6, atol=1e-7)


def test_fft_pair_matches_numpy(gpu):
    g = pkg("gpu")
    D = 48
    x = np.random.default_rng(0).standard_normal((D, D, D)).astype(np.float32)
    spec = g.r2c(_dev(x)).cpu().numpy()
    want = np.fft.rfftn(x.astype(np.float64))
    assert np.abs(spec - want).max() < 2e-6 * np.abs(want).max()
    back = g.c2r(_dev(spec)).cpu().numpy()
    assert np.abs(back / D ** 3 - x).max() < 2e-6 * np.abs(x).max()


def test_fcoll_sum_and_threshold(gpu):
    g = pkg("gpu")
    D = 32
    rng = np.random.default_rng(5)
    sm = (rng.standard_normal((D, D, D)) * 0.8).astype(np.float32)
    deltac, denom, zeta, fc_mean = 1.686, np.float32(1.3), 30.0, np.float32(0.02)
    fc = bubbles.fcoll_f32(sm, deltac, denom)
    s = g.fcoll_sum(_dev(sm), deltac, denom)
    assert abs(float(s.item()) / fc.astype(np.float64).sum() - 1) < 1e-6
    # linear threshold, non-final then final
    ion = torch.zeros((D, D, D), device="cuda")
    g.fcoll_threshold(ion, _dev(sm), deltac, denom, zeta)
    hit = fc.astype(np.float64) >= 1 / zeta
    assert np.array_equal(ion.cpu().numpy() == 1, hit)
    ion2 = torch.zeros((D, D, D), device="cuda")
    g.fcoll_threshold(ion2, _dev(sm), deltac, denom, zeta, final_step=True)
    want = np.where(hit, 1.0, (fc.astype(np.float64) * zeta)).astype(np.float32)
    assert np.allclose(ion2.cpu().numpy(), want, rtol=1e-5, atol=1e-7)
    # normalised (nonlinear input): f * float32(fc_mean / mean)
    ion3 = torch.zeros((D, D, D), device="cuda")
    g.fcoll_threshold(ion3, _dev(sm), deltac, denom, zeta, dev_sum=s, fc_mean=fc_mean, final_step=True)
    scale = np.float32(float(fc_mean) / (fc.astype(np.float64).sum() / D ** 3))
    fcn = (fc * scale).astype(np.float32)
    hitn = fcn.astype(np.float64) >= 1 / zeta
    got = ion3.cpu().numpy()
    band = np.abs(zeta * fcn.astype(np.float64) - 1) < 1e-5
    assert np.array_equal((got == 1)[~band], hitn[~band])
    assert np.allclose(got[~hitn & ~band], (fcn.astype(np.float64) * zeta)[~hitn & ~band], rtol=1e-5)


@pytest.mark.parametrize("periodic", [False, True])
def test_sphere_paint_matches_oracle(gpu, periodic):
    g = pkg("gpu")
    D = 24
    rng = np.random.default_rng(7)
    sm = (rng.standard_normal((D, D, D)) * 0.45).astype(np.float32)
    # zeta = 1.6: centres where delta >~ 1.34, about 3 sigma (a few dozen cells)
    deltac, denom, zeta = 1.686, np.float32(1.0), 1.6
    fc = bubbles.fcoll_f32(sm, deltac, denom)
    centers = fc.astype(np.float64) >= 1 / zeta
    assert 0 < centers.sum() < 80
    assert np.min(np.abs(fc.astype(np.float64) * zeta - 1.0)) > 1e-6  # no cell on the threshold
    for R in (1.5, 3.7, 9.2):
        ion = torch.zeros((D, D, D), device="cuda")
        g.update_sphere(ion, _dev(sm), deltac, denom, zeta, R, periodic)
        want = bubbles.paint_spheres(np.zeros((D, D, D), np.float32), centers, np.float32(R), periodic)
        assert np.array_equal(ion.cpu().numpy(), want), R


def _within_grf_tolerance(got, want):
    """The GPU draws the normals and amplitudes in float32 from the same
    Philox words (exact uniforms and angles), the oracle in FP64: each mode
    agrees to a few float32 ulp of its modulus, and both halves of the
    Hermitian pairs and the real self-conjugate modes are where they belong."""
    w = want.astype(np.complex128)
    diff = np.abs(got.astype(np.complex128) - w)
    assert np.all(diff <= 2e-6 * np.abs(w) + 1e-30), float(np.max(diff / np.maximum(np.abs(w), 1e-30)))
    assert np.array_equal(got == 0, want == 0)


def test_grf_matches_oracle(gpu):
    g = pkg("gpu")
    sig = pkg("cosmo_files.sigmas")
    for D, L, seed, box in ((16, 40.0, 3, 0), (64, 100.0, 0, 5)):
        amp = sig.PowerLawSpectrum(D, L).amp
        got = g.grf_init(D, L, amp, seed, box).cpu().numpy()
        want = grf.grf_half_spectrum(D, L, seed, box)
        _within_grf_tolerance(got, want)


def test_grf_table_matches_oracle(gpu):
    """cm21_grf_init_table with the Planck13 table of initialize.py:27-31
    against the oracle's table interpolation (same Philox stream)."""
    g = pkg("gpu")
    pl = pkg("cosmo_files.planck")
    for D, L, seed, box in ((16, 40.0, 3, 0), (64, 128.0, 0, 5)):
        K, Pk = pl.planck_spectrum().init_table(D, L, 512)
        got = g.grf_init_table(D, L, K, Pk, seed, box).cpu().numpy()
        want = grf.grf_half_spectrum(D, L, seed, box, pk_table=(K, Pk))
        _within_grf_tolerance(got, want)
    # a short table: modes beyond its last K get no power
    K = np.array([0.05, 0.2, 0.5], np.float32)
    Pk = np.array([10.0, 5.0, 1.0], np.float32)
    got = g.grf_init_table(16, 40.0, K, Pk, 1).cpu().numpy()
    want = grf.grf_half_spectrum(16, 40.0, 1, pk_table=(K, Pk))
    _within_grf_tolerance(got, want)
    assert np.any(got == 0) and np.any(got != 0)
    # a table spaced linearly in k: the search starts from a log-uniform
    # guess and walks further to the same index
    K = np.linspace(0.01, 3.0, 300).astype(np.float32)
    Pk = (100.0 / (1.0 + (K / 0.3) ** 2)).astype(np.float32)
    got = g.grf_init_table(16, 40.0, K, Pk, 2).cpu().numpy()
    want = grf.grf_half_spectrum(16, 40.0, 2, pk_table=(K, Pk))
    _within_grf_tolerance(got, want)


def test_init_pipeline_planck_table_matches_oracle(gpu):
    init = pkg("Programs.initialize")
    D, sw, L, seed = 64, 32, 128.0, 4
    out = init.init(DIM=D, HII_DIM=sw, BOX_LEN=L, seed=seed, save=False)
    want = grf.initialize(D, sw, L, seed, pk_table=out["pk_table"])
    for key in ("deltax", "smoothed"):
        a = out[key].cpu().numpy()
        assert np.abs(a - want[key]).max() < 1e-5 * np.abs(want[key]).max(), key
    for c in range(3):
        a = out["v"][c].cpu().numpy()
        assert np.abs(a - want["v"][c]).max() < 1e-5 * np.abs(want["v"][c]).max(), c


def test_subsample_and_velocity(gpu):
    g = pkg("gpu")
    D, sw, L = 32, 16, 60.0
    big = np.random.default_rng(1).standard_normal((D, D, D)).astype(np.float32)
    assert np.array_equal(g.subsample(_dev(big), sw).cpu().numpy(), grf.subsample(big, sw))
    # non-integer pixel factor
    assert np.array_equal(g.subsample(_dev(big), 12).cpu().numpy(), grf.subsample(big, 12))
    s = grf.grf_half_spectrum(D, L, seed=2)
    R = grf.smoothing_radius(L, sw)
    km = bubbles.kmag_half_f32(D, bubbles.delta_k_f32(L), fastest_first=False)
    w0, cut, w2 = bubbles.window_f32(km, R, 0, True)
    for comp in range(3):
        got = g.velocity(_dev(s), L, comp, 0, R).cpu().numpy()
        v = (s.astype(np.complex128) * grf.velocity_multiplier(D, L, comp)).astype(np.complex64)
        want = bubbles.apply_window_f32(v, w0, cut, w2)
        assert np.allclose(got, want, rtol=1e-5, atol=1e-6 * np.abs(want).max())


def test_init_pipeline_matches_oracle(gpu):
    init = pkg("Programs.initialize")
    D, sw, L, seed = 64, 32, 100.0, 9
    out = init.init(DIM=D, HII_DIM=sw, BOX_LEN=L, seed=seed, pspec="powerlaw", save=False)
    want = grf.initialize(D, sw, L, seed)
    for key in ("deltax", "smoothed"):
        a = out[key].cpu().numpy()
        assert np.abs(a - want[key]).max() < 1e-5 * np.abs(want[key]).max(), key
    for c in range(3):
        a = out["v"][c].cpu().numpy()
        assert np.abs(a - want["v"][c]).max() < 1e-5 * np.abs(want["v"][c]).max(), c


@pytest.mark.parametrize("DIM,sw,amp", [(32, 16, 10.0), (64, 32, 40.0), (64, 64, 3.0), (48, 24, 8.0)])
def test_move_mass_matches_oracle(gpu, DIM, sw, amp):
    g = pkg("gpu")
    rng = np.random.default_rng(DIM + sw)
    L = 2.0 * DIM
    delta = rng.standard_normal((DIM,) * 3).astype(np.float32)
    v = [(rng.standard_normal((sw,) * 3) * amp).astype(np.float32) for _ in range(3)]
    s = perturb.displacement_scale(0.12, 0.0036, L)
    dv = [_dev(x) for x in v]
    # zero density contrast: every deposit is exactly 1, so float sums are
    # exact and the NGP cells are checked bit for bit
    z = np.zeros_like(delta)
    got = g.move_mass(_dev(z), *dv, s, 0.0036, "ngp").cpu().numpy()
    assert np.array_equal(got, perturb.move_mass_ngp(z, *v, 0.12, 0.0036, L))
    got = g.move_mass(_dev(delta), *dv, s, 0.0036, "ngp").cpu().numpy()
    assert np.allclose(got, perturb.move_mass_ngp(delta, *v, 0.12, 0.0036, L), rtol=1e-5, atol=2e-6)
    if DIM % 16 == 0:
        got = g.move_mass(_dev(delta), *dv, s, 0.0036, "cic").cpu().numpy()
        assert np.allclose(got, perturb.move_mass_cic(delta, *v, 0.12, 0.0036, L), rtol=1e-5, atol=2e-6)


def test_delta_T_and_power_spectrum(gpu):
    g = pkg("gpu")
    D, L, z = 32, 64.0, 8.0
    rng = np.random.default_rng(4)
    xH = (rng.random((D, D, D)) > 0.3).astype(np.float32)
    d = (rng.standard_normal((D, D, D)) * 0.3).astype(np.float32)
    c = deltat.tb_constant(z)
    dT, s = g.delta_T(_dev(xH), _dev(d), c)
    want = deltat.delta_T_f32(xH, d, z)
    assert np.array_equal(dT.cpu().numpy(), want)
    assert abs(s.item() / want.astype(np.float64).sum() - 1) < 1e-12
    edges = deltat.log_edges(L, D, 20)
    ps, ks, cnt = g.power_spectrum(dT, L, edges, field_sum=s)
    k_ave, ps_ave, cnt_want, (ps_want, ks_want) = deltat.power_spectrum(want, L, edges)
    assert np.array_equal(cnt.cpu().numpy(), cnt_want)
    assert np.allclose(ks.cpu().numpy(), ks_want, rtol=1e-12)
    assert np.allclose(ps.cpu().numpy(), ps_want, rtol=1e-4)
    # deterministic
    ps2, _, _ = g.power_spectrum(dT, L, edges, field_sum=s)
    assert torch.equal(ps, ps2)
