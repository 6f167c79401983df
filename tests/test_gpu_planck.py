"""conv_bubbles in the Planck-cosmology mode (no deltac / growth / sigma_min2
given: the barrier from Z, INIT_PARAMS.ION_Tvir_MIN and ZETA, S(R) of the
Planck13 spectrum) equals the call with the same scalars passed explicitly,
and matches the FP32 oracle outside the threshold band."""
import numpy as np
import pytest
from scipy.special import erfc

from conftest import pkg
from oracle import bubbles, grf, sigma

pytestmark = pytest.mark.gpu


def test_planck_mode_matches_explicit_and_oracle(gpu):
    pl = pkg("cosmo_files.planck")
    fb = pkg("Programs.find_HII_bubbles")
    P = pkg("Parameter_files.INIT_PARAMS")
    D, L, z = 64, 100.0, 9.0
    delta = grf.c2r(grf.grf_half_spectrum(D, L, seed=21), D).astype(np.float32)
    pd = {"BoxSize": L, "HIIdim": D, "Z": z}
    print(f"planck mode D={D} z={z} zeta={P.ZETA}", flush=True)
    got = fb.conv_bubbles(delta, pd, LE=True, quiet=True)
    b = pl.barrier(z)
    exp = fb.conv_bubbles(delta, pd, LE=True, quiet=True, deltac=b["deltac"], sigma_min2=b["sigma_min2"],
                          sig0=b["sig0"])
    assert np.array_equal(got, exp)
    radii, _ = sigma.radius_schedule(L, D)
    den = sigma.denominators(radii, b["sigma_min2"], b["sig0"])
    fc = np.float32(erfc(b["deltac"] / np.sqrt(2 * b["sigma_min2"])))
    gi = np.float32(b["deltac"] / 1.686)
    want = bubbles.find_bubbles_f32(delta, L, radii, den, P.ZETA, b["deltac"], LE=True, growth_inv=gi, fc_mean=fc)
    _, margin = bubbles.find_bubbles_f64(delta, L, radii, den, P.ZETA, b["deltac"], LE=True,
                                         growth_inv=float(b["deltac"]) / 1.686, fc_mean=float(fc))
    ok = margin > 1e-5
    assert np.mean(ok) > 0.99
    assert np.array_equal((got == 1)[ok], (want == 1)[ok])
    part = ok & (want < 1) & (got < 1)
    assert np.allclose(got[part], want[part], rtol=1e-4, atol=1e-7)
    # a partly ionized box: flagged cells and painted partial values both occur
    x_hi = 1.0 - float(got.mean())
    print(f"x_HI {x_hi:.4f}, flagged {float((got == 1).mean()):.3f}", flush=True)
    assert 0.02 < x_hi < 0.5 and 0.5 < float((got == 1).mean()) < 0.99


def test_batched_planck_mode_matches_single_box(gpu):
    """Boxes of a batch given only their z take the Planck13 barrier and S(R)
    and agree with conv_bubbles' Planck mode box by box."""
    fb = pkg("Programs.find_HII_bubbles")
    bm = pkg("Programs.find_HII_bubbles_batched")
    D, L = 64, 100.0
    zs = [9.0, 9.5]
    deltas = [grf.c2r(grf.grf_half_spectrum(D, L, seed=30 + i), D).astype(np.float32) for i in range(len(zs))]
    specs = [{"D": D, "box_len": L, "zeta": 40.0, "z": z} for z in zs]
    bb = bm.BatchedBubbles(specs, fil=1, LE=True, graph=False)
    bb.set_deltas(deltas)
    bb.run()
    for i, z in enumerate(zs):
        got = bb.ionized(i).cpu().numpy()
        single = fb.conv_bubbles(deltas[i], {"BoxSize": L, "HIIdim": D, "Z": z}, LE=True, quiet=True, zeta=40.0)
        print(f"box {i} z={z}: x_HI {1 - got.mean():.4f}", flush=True)
        assert np.mean((got == 1) != (single == 1)) <= 1e-5
        both = (got < 1) & (single < 1)
        assert np.allclose(got[both], single[both], rtol=0, atol=1e-5)


def test_velocities_at_z_are_growth_rate_times_initial(gpu):
    """perturb_field.velocities_at (perturb_field.py:111-126) from the z = 0
    density equals (dD/dt)/D times initialize's velocity boxes (the same
    filtered, subsampled psi), box for box."""
    init = pkg("Programs.initialize")
    pf = pkg("Programs.perturb_field")
    pl = pkg("cosmo_files.planck")
    DIM, HII, L, z = 64, 32, 50.0, 9.0
    ics = init.init(DIM=DIM, HII_DIM=HII, BOX_LEN=L, seed=5, save=False)
    f = pl.ddt_growth_over_growth(z)
    vs = pf.velocities_at(z, ics["deltax"], HII_DIM=HII, box_len=L)
    # against the oracle's initialize velocities on the same spectrum table
    from oracle import grf
    want = grf.initialize(DIM, HII, L, 5, pk_table=ics["pk_table"])["v"]
    for axis, (v, w) in enumerate(zip(vs, want)):
        ref = w.astype(np.float64) * f
        err = float(np.abs(v.double().cpu().numpy() - ref).max())
        assert err <= 1e-4 * float(np.abs(ref).max()), axis
    for axis, (v, v0) in enumerate(zip(vs, ics["v"])):
        ref = (v0.double() * f).cpu()
        err = float((v.double().cpu() - ref).abs().max())
        scale = float(ref.abs().max())
        print(f"axis {axis}: max err {err:.3g} of {scale:.3g}", flush=True)
        assert err <= 1e-4 * scale
