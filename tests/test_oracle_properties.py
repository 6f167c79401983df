"""Property-based checks of the excursion-set oracle (CPU, hypothesis).

Invariants of the algorithm itself (the radius loop of find_HII_bubbles.py:
113-193 with the kernels of find_bubbles.cu:43-71,142-158), none of which
restates its formulas, on random Gaussian boxes and parameters:

* raising zeta never lowers the ionized field anywhere (the threshold is
  f_coll >= 1/zeta, the final paint zeta f_coll): a reversed comparison or a
  swapped threshold fails;
* adding a positive constant to the density never lowers it (f_coll =
  erfc((deltac - delta_R)/denom) grows with delta_R, and the window keeps the
  mean): a sign slip in the erfc argument fails;
* the field commutes with periodic shifts and with permutations of the axes
  (the window depends on |k| only): an axis mix-up in the half spectrum, a
  wrong k for the halved axis or an edge effect fails.

Cells within 1e-9 of the threshold (the oracle's margin) are left out of the
flag comparisons where two runs round differently.  The monotonicity checks
only count partly ionized boxes (a box ionized everywhere satisfies them
trivially), so an oracle that ionizes every cell leaves no valid example and
fails.
"""
import numpy as np
from hypothesis import HealthCheck, assume, given, settings
from hypothesis import strategies as st

from oracle import bubbles, grf, sigma

D, L = 16, 30.0
RADII, _ = sigma.radius_schedule(L, D)
BAND = 1e-9

COMMON = settings(max_examples=12, deadline=None, derandomize=True, suppress_health_check=[HealthCheck.too_slow])
SEEDS = st.integers(0, 100_000)
SMIN2 = st.floats(0.5, 1.5)


def _case(seed, smin2):
    delta = grf.c2r(grf.grf_half_spectrum(D, L, seed=seed), D).astype(np.float32)
    den = sigma.denominators(RADII, smin2, lambda R: sigma.sig0_discrete(R, D, L))
    return delta, den


def _f64(delta, den, zeta, deltac=1.686):
    return bubbles.find_bubbles_f64(np.ascontiguousarray(delta), L, RADII, den, zeta, deltac)


@COMMON
@given(seed=SEEDS, smin2=SMIN2, z1=st.floats(2.0, 60.0), z2=st.floats(2.0, 60.0))
def test_ionized_field_monotone_in_zeta(seed, smin2, z1, z2):
    lo_z, hi_z = min(z1, z2), max(z1, z2)
    delta, den = _case(seed, smin2)
    # the float32 restatement (the parity target) and the float64 one
    lo = bubbles.find_bubbles_f32(delta, L, RADII, den, lo_z, 1.686)
    assume(0.0 < (lo == 1).mean() < 1.0)
    hi = bubbles.find_bubbles_f32(delta, L, RADII, den, hi_z, 1.686)
    assert np.all(hi >= lo)
    lo64, hi64 = _f64(delta, den, lo_z)[0], _f64(delta, den, hi_z)[0]
    assert np.all(hi64 >= lo64)


@COMMON
@given(seed=SEEDS, smin2=SMIN2, zeta=st.floats(5.0, 60.0), c=st.floats(0.01, 0.5))
def test_ionized_field_monotone_in_density(seed, smin2, zeta, c):
    delta, den = _case(seed, smin2)
    base, m0 = _f64(delta, den, zeta)
    assume(0.0 < (base == 1).mean() < 1.0)
    up, m1 = _f64(delta + np.float32(c), den, zeta)
    ok = (m0 > BAND) & (m1 > BAND)
    assert np.all((up == 1)[ok & (base == 1)])
    part = ok & (base < 1) & (up < 1)
    assert np.all(up[part] >= base[part] - 1e-12)
    assert (up == 1).sum() >= (base == 1).sum()


@COMMON
@given(seed=SEEDS, smin2=SMIN2, zeta=st.floats(5.0, 60.0),
       shift=st.tuples(st.integers(0, D - 1), st.integers(0, D - 1), st.integers(0, D - 1)))
def test_commutes_with_periodic_shift(seed, smin2, zeta, shift):
    delta, den = _case(seed, smin2)
    ion, m = _f64(delta, den, zeta)
    ion_s, m_s = _f64(np.roll(delta, shift, axis=(0, 1, 2)), den, zeta)
    want, mw = np.roll(ion, shift, axis=(0, 1, 2)), np.roll(m, shift, axis=(0, 1, 2))
    ok = (mw > BAND) & (m_s > BAND)
    assert np.array_equal((ion_s == 1)[ok], (want == 1)[ok])
    assert np.allclose(ion_s[ok], want[ok], rtol=0, atol=1e-9)


@COMMON
@given(seed=SEEDS, smin2=SMIN2, zeta=st.floats(5.0, 60.0), perm=st.permutations([0, 1, 2]))
def test_commutes_with_axis_permutation(seed, smin2, zeta, perm):
    delta, den = _case(seed, smin2)
    ion, m = _f64(delta, den, zeta)
    ion_p, m_p = _f64(np.transpose(delta, perm), den, zeta)
    want, mw = np.transpose(ion, perm), np.transpose(m, perm)
    ok = (mw > BAND) & (m_p > BAND)
    assert np.array_equal((ion_p == 1)[ok], (want == 1)[ok])
    assert np.allclose(ion_p[ok], want[ok], rtol=0, atol=1e-9)
    # the float32 restatement too (its transforms round differently per axis order)
    f = bubbles.find_bubbles_f32(np.ascontiguousarray(np.transpose(delta, perm)), L, RADII, den, zeta, 1.686)
    g = np.transpose(bubbles.find_bubbles_f32(delta, L, RADII, den, zeta, 1.686), perm)
    ok32 = (mw > 1e-5) & (m_p > 1e-5)
    assert np.array_equal((f == 1)[ok32], (g == 1)[ok32])


# ---------------------------------------------------------------- P21(k)
# deltat.power_spectrum (delta_T.py:80-98 with the conventions kept from
# delta_T.cu:5-29): the binned power of delta_T/<delta_T> - 1 does not change
# under a global rescaling of the field (the dimensional one scales as its
# square), under periodic shifts and under permutations of the axes, which
# moves the halved r2c axis (a wrong weight of its 0 and Nyquist planes
# fails).

from oracle import deltat  # noqa: E402

PD, PL = 16, 50.0
EDGES = deltat.log_edges(PL, PD, 12)


def _field(seed):
    rng = np.random.default_rng(seed)
    return 1.0 + 0.3 * rng.standard_normal((PD, PD, PD))


def _ps(f, dimensional=False):
    return deltat.power_spectrum(f, PL, EDGES, dimensional=dimensional)


@COMMON
@given(seed=SEEDS, a=st.floats(0.01, 100.0))
def test_power_spectrum_scaling(seed, a):
    f = _field(seed)
    k0, p0, c0, _ = _ps(f)
    k1, p1, c1, _ = _ps(a * f)
    assert np.array_equal(c0, c1) and np.array_equal(k0, k1)
    assert np.allclose(p1, p0, rtol=1e-10, atol=0)
    pd0, pd1 = _ps(f, True)[1], _ps(a * f, True)[1]
    assert np.allclose(pd1, a * a * pd0, rtol=1e-10, atol=0)


@COMMON
@given(seed=SEEDS, perm=st.permutations([0, 1, 2]),
       shift=st.tuples(st.integers(0, PD - 1), st.integers(0, PD - 1), st.integers(0, PD - 1)))
def test_power_spectrum_shift_and_axis_invariance(seed, perm, shift):
    f = _field(seed)
    k0, p0, c0, _ = _ps(f)
    for g in (np.roll(f, shift, axis=(0, 1, 2)), np.ascontiguousarray(np.transpose(f, perm))):
        k1, p1, c1, _ = _ps(g)
        assert np.array_equal(c0, c1)
        assert np.allclose(k1, k0, rtol=1e-12, atol=0)
        assert np.allclose(p1, p0, rtol=1e-9, atol=0)


# ---------------------------------------------------------------- mass move
# perturb.move_mass_ngp / _cic (perturb_field.py:65-108, perturb_field.cu:
# 90-133): permuting the axes of the density and of the displacement fields
# (component a of the new field is the permuted component perm[a]) permutes
# the moved box; mixing up displacement components or the axis a component's
# lattice term is laid along fails.

from oracle import perturb  # noqa: E402


@COMMON
@given(seed=SEEDS, perm=st.permutations([0, 1, 2]), sw=st.sampled_from([4, 8]), scheme=st.sampled_from(["ngp", "cic"]))
def test_mass_move_commutes_with_axis_permutation(seed, perm, sw, scheme):
    rng = np.random.default_rng(seed)
    DIM = 8
    delta = rng.standard_normal((DIM, DIM, DIM)).astype(np.float32)
    v = [(rng.standard_normal((sw, sw, sw)) * 3.0).astype(np.float32) for _ in range(3)]
    fn = perturb.move_mass_ngp if scheme == "ngp" else perturb.move_mass_cic
    args = (0.5, 0.1, 10.0)  # growth, initial growth, box length: displacements of about a cell
    base = fn(delta, *v, *args)
    pd = np.ascontiguousarray(np.transpose(delta, perm))
    pv = [np.ascontiguousarray(np.transpose(v[perm[a]], perm)) for a in range(3)]
    got = fn(pd, *pv, *args)
    assert np.allclose(got, np.transpose(base, perm), rtol=1e-6, atol=1e-6)
