"""Pins the nonlinear-input branch (LE=False) of the excursion-set oracle.

The reference linearly extrapolates a nonlinear density box to z = 0
(find_HII_bubbles.py:58,84-85: deltax *= deltac/1.686) and, at every radius,
rescales f_coll so that its box mean equals the mean collapsed fraction
fc_mean (:158-161,171-174: fcoll *= fc_mean_ps/fcollmean).  None of the checks
below retypes those lines; each follows from what they imply:

* a constant box stays constant under every window (the DC mode has window
  1), so after the rescaling every cell holds fc_mean whatever the density,
  barrier, growth or radius: the box ionizes exactly when zeta*fc_mean >= 1,
  otherwise the final paint is zeta*fc_mean everywhere;
* on any box the rescaled f_coll has box mean fc_mean at every radius;
* the growth enters only through the input: the loop on delta with growth
  1/g equals the loop on g-scaled delta with growth 1;
* the rescaled f_coll at a radius is proportional, cell by cell, to the
  unscaled one of the same input (a single scalar per radius);
* raising zeta never lowers the ionized field (hypothesis, as for LE=True).

``test_planted_mistakes_fail`` plants an inverted scale, a sum in place of
the mean and the growth applied twice into the oracle and shows that each
makes at least one of these checks fail.
"""
import numpy as np
import pytest
from hypothesis import HealthCheck, assume, given, settings
from hypothesis import strategies as st

from oracle import bubbles, grf, sigma

D, L = 16, 30.0
RADII, _ = sigma.radius_schedule(L, D)


def _denoms(smin2):
    return sigma.denominators(RADII, smin2, lambda R: sigma.sig0_discrete(R, D, L))


def _grf(seed):
    return grf.c2r(grf.grf_half_spectrum(D, L, seed=seed), D).astype(np.float32)


def _run(delta, den, zeta, deltac, g_inv, fc_mean, trace=False):
    return bubbles.find_bubbles_f32(delta, L, RADII, den, zeta, deltac, LE=False, growth_inv=np.float32(g_inv),
                                    fc_mean=np.float32(fc_mean), return_trace=trace)


# ---------------------------------------------------------------- checks

def check_constant_box():
    den = _denoms(0.6)
    for c, g_inv, deltac in ((0.2, 2.0, 1.686), (-0.5, 0.7, 2.4), (1.3, 4.0, 3.0)):
        delta = np.full((D, D, D), c, np.float32)
        for fc_mean, zeta in ((0.02, 30.0), (0.07, 30.0), (0.3, 2.0), (0.3, 4.0)):
            got = _run(delta, den, zeta, deltac, g_inv, fc_mean)
            target = zeta * fc_mean
            assert abs(target - 1.0) > 1e-3
            if target >= 1.0:
                assert np.all(got == 1.0), (c, g_inv, fc_mean, zeta)
            else:
                assert np.allclose(got, target, rtol=2e-6, atol=0), (c, g_inv, fc_mean, zeta, got.min(), got.max())


def check_normalised_mean():
    for seed, smin2, g_inv, deltac, fc_mean in ((1, 0.9, 1.0 / 0.3, 1.686 / 0.3, 0.05),
                                                (2, 1.2, 2.0, 3.372, 0.011), (3, 0.8, 1.25, 2.1, 0.2)):
        den = _denoms(smin2)
        _, tr = _run(_grf(seed), den, 30.0, deltac, g_inv, fc_mean, trace=True)
        assert len(tr) == len(RADII)
        for t in tr:
            m = float(np.mean(t["fcoll"].astype(np.float64)))
            assert abs(m / float(np.float32(fc_mean)) - 1.0) < 2e-6, (seed, float(t["R"]), m, fc_mean)


def check_growth_scaling():
    n_partial = 0
    for seed, g in ((4, 0.25), (5, 0.5), (6, 0.12)):
        delta = _grf(seed)
        den = _denoms(1.0)
        deltac = 1.686 / g
        fc_mean = 0.03
        a = _run(delta, den, 30.0, deltac, 1.0 / g, fc_mean)
        scaled = (delta * np.float32(1.0 / g)).astype(np.float32)
        b = _run(scaled, den, 30.0, deltac, 1.0, fc_mean)
        assert np.array_equal(a, b), seed
        n_partial += int(0.0 < (a == 1).mean() < 1.0)
    assert n_partial >= 2  # the comparison must see ionized and neutral cells


def check_single_scalar_per_radius():
    delta, den = _grf(7), _denoms(1.0)
    g_inv, deltac, fc_mean = 3.0, 1.686 * 3.0, 0.04
    _, tr_nl = _run(delta, den, 30.0, deltac, g_inv, fc_mean, trace=True)
    scaled = (delta * np.float32(g_inv)).astype(np.float32)
    _, tr_lin = bubbles.find_bubbles_f32(scaled, L, RADII, den, 30.0, deltac, LE=True, return_trace=True)
    for a, b in zip(tr_nl, tr_lin):
        f_nl, f_lin = a["fcoll"].astype(np.float64), b["fcoll"].astype(np.float64)
        live = f_lin > 1e-30
        ratio = f_nl[live] / f_lin[live]
        assert ratio.max() / ratio.min() - 1.0 < 1e-6, float(a["R"])


CHECKS = [check_constant_box, check_normalised_mean, check_growth_scaling, check_single_scalar_per_radius]


@pytest.mark.parametrize("check", CHECKS, ids=lambda c: c.__name__)
def test_nonlinear_branch(check):
    check()


def test_nonlinear_golden_vector_still_holds():
    # the committed D=32 nonlinear vector was written by this oracle; with the
    # checks above pinning the branch it stays a regression vector
    import os
    from conftest import GOLDEN
    d = np.load(os.path.join(GOLDEN, "bubbles_D32_nl.npz"))
    got = bubbles.find_bubbles_f32(d["delta"], float(d["box_len"]), d["radii"], d["denoms"], float(d["zeta"]),
                                   float(d["deltac"]), LE=False, growth_inv=np.float32(float(d["deltac"]) / 1.686),
                                   fc_mean=d["fc_mean"])
    assert np.array_equal(got, d["ionized"])
    assert 0.0 < (got == 1).mean() < 1.0


def test_zero_box_mean_is_ieee():
    # f_coll underflows to 0 in every cell: the reference divides fc_mean by a
    # zero box mean (an infinite scale), so 0*inf paints NaN at the last radius
    delta = np.zeros((D, D, D), np.float32)
    with np.errstate(all="ignore"):
        got = _run(delta, _denoms(0.6), 30.0, 60.0, 1.0, 0.01)
        got64 = bubbles.find_bubbles_f64(delta, L, RADII, _denoms(0.6), 30.0, 60.0, LE=False, growth_inv=1.0,
                                         fc_mean=0.01)[0]
    assert np.all(np.isnan(got)) and np.all(np.isnan(got64))


# ---------------------------------------------------------------- planted mistakes

def _inverted_scale(fc, fc_mean):
    mean = float(np.sum(np.asarray(fc, np.float64))) / np.asarray(fc).size
    return np.float32(mean / float(np.float32(fc_mean)))


def _sum_not_mean(fc, fc_mean):
    return np.float32(float(np.float32(fc_mean)) / float(np.sum(np.asarray(fc, np.float64))))


def _growth_twice(d, growth_inv):
    g = np.float32(growth_inv)
    return ((np.asarray(d, np.float32) * g).astype(np.float32) * g).astype(np.float32)


PLANTS = {
    "inverted_scale": ("box_mean_scale_f32", _inverted_scale),
    "sum_not_mean": ("box_mean_scale_f32", _sum_not_mean),
    "growth_twice": ("nonlinear_input_f32", _growth_twice),
}


@pytest.mark.parametrize("plant", sorted(PLANTS))
def test_planted_mistakes_fail(plant, monkeypatch):
    name, fn = PLANTS[plant]
    monkeypatch.setattr(bubbles, name, fn)
    failed = []
    for check in CHECKS:
        try:
            check()
        except AssertionError:
            failed.append(check.__name__)
    assert failed, f"planted {plant} passed every nonlinear check"


# ---------------------------------------------------------------- monotone in zeta

@settings(max_examples=10, deadline=None, derandomize=True, suppress_health_check=[HealthCheck.too_slow])
@given(seed=st.integers(0, 100_000), smin2=st.floats(0.7, 1.5), g=st.floats(0.25, 1.0),
       fc_mean=st.floats(0.005, 0.2), z1=st.floats(2.0, 80.0), z2=st.floats(2.0, 80.0))
def test_nonlinear_monotone_in_zeta(seed, smin2, g, fc_mean, z1, z2):
    lo_z, hi_z = min(z1, z2), max(z1, z2)
    delta, den = _grf(seed), _denoms(smin2)
    lo = _run(delta, den, lo_z, 1.686 / g, 1.0 / g, fc_mean)
    assume(0.0 < (lo == 1).mean() < 1.0)
    hi = _run(delta, den, hi_z, 1.686 / g, 1.0 / g, fc_mean)
    assert np.all(hi >= lo)
    lo64 = bubbles.find_bubbles_f64(delta, L, RADII, den, lo_z, 1.686 / g, LE=False, growth_inv=1.0 / g,
                                    fc_mean=fc_mean)[0]
    hi64 = bubbles.find_bubbles_f64(delta, L, RADII, den, hi_z, 1.686 / g, LE=False, growth_inv=1.0 / g,
                                    fc_mean=fc_mean)[0]
    assert np.all(hi64 >= lo64)
