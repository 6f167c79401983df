"""The redshift scroll (Programs/Drives.py): one initialization, the density
at each redshift, and one batched find_HII_bubbles call for all redshifts,
checked against the stage functions called one by one."""
import numpy as np
import pytest
import torch

from conftest import pkg

pytestmark = pytest.mark.gpu


def test_zscroll_matches_stage_by_stage(gpu):
    drv = pkg("Programs.Drives")
    init = pkg("Programs.initialize")
    pert = pkg("Programs.perturb_field")
    fb = pkg("Programs.find_HII_bubbles")
    bm = pkg("Programs.find_HII_bubbles_batched")
    DIM, D, L = 128, 64, 100.0
    # sigma_min^2 in z=0 units such that the barrier 1.686/D(z) is crossed in part of the box
    out = drv.drive_zscroll_noTs(12.0, 6.0, -2.0, DIM=DIM, HII_DIM=D, BOX_LEN=L, seed=3, zeta=30.0, sigma_min2=40.0,
                                 quiet=True)
    assert list(out["z"]) == [12.0, 10.0, 8.0]
    ics = init.init(DIM=DIM, HII_DIM=D, BOX_LEN=L, seed=3, pspec="powerlaw", save=False)
    d_init = bm.linear_growth(300.0)
    for i, z in enumerate(out["z"]):
        upd = pert.evolve_zeldovich(z, ics["deltax"], ics["v"], bm.linear_growth(z), d_init, L, "cic", save=False)
        # the CIC deposit uses float atomics: equal up to the order of additions
        assert torch.allclose(upd, out["density"][i], rtol=0, atol=1e-5)
        single = fb.conv_bubbles(out["density"][i], {"BoxSize": L, "HIIdim": D}, LE=False, quiet=True,
                                 deltac=1.686 / bm.linear_growth(z), sigma_min2=40.0, zeta=30.0).cpu().numpy()
        got = out["ionized"][i].cpu().numpy()
        print(f"z={z}: x_HI {out['x_HI'][i]:.4f}", flush=True)
        assert np.mean((got == 1) != (single == 1)) <= 1e-5
        both = (got < 1) & (single < 1)
        assert np.allclose(got[both], single[both], rtol=0, atol=1e-5)
        K, k_ave, ps = out["P21"][i]
        assert len(ps) == 40 and np.all(np.isfinite(ps))
    # later redshifts are more ionized
    assert out["x_HI"][0] >= out["x_HI"][1] >= out["x_HI"][2]
    # the lowest redshift against the FP64 oracle on the scroll's own density
    from oracle import bubbles, sigma
    from scipy.special import erfc
    sp = pkg("cosmo_files.sigmas").PowerLawSpectrum(D, L)
    z, dens = float(out["z"][-1]), out["density"][-1].cpu().numpy()
    dc = 1.686 / bm.linear_growth(z)
    radii, _ = sigma.radius_schedule(L, D)
    den = sigma.denominators(radii, 40.0, sp.sig0_discrete)
    fcm = float(np.float32(erfc(dc / np.sqrt(80.0))))
    want, margin = bubbles.find_bubbles_f64(dens, L, radii, den, 30.0, dc, LE=False,
                                            growth_inv=float(np.float32(dc / 1.686)), fc_mean=fcm)
    got = out["ionized"][-1].cpu().numpy()
    ok = margin > 1e-5
    assert np.array_equal((got == 1)[ok], (want == 1)[ok])
    part = ok & (got < 1) & (want < 1)
    assert np.allclose(got[part], want[part], rtol=1e-4, atol=1e-7)


def test_zscroll_planck_mode(gpu):
    """Without sigma_min2 the scroll follows the Planck13 cosmology (D(z) and the
    per-redshift barrier of cosmo_files.planck); each redshift's box equals
    conv_bubbles' Planck mode on the same density."""
    drv = pkg("Programs.Drives")
    fb = pkg("Programs.find_HII_bubbles")
    DIM, D, L = 128, 64, 100.0
    out = drv.drive_zscroll_noTs(12.0, 9.5, -1.0, DIM=DIM, HII_DIM=D, BOX_LEN=L, seed=4, quiet=True)
    assert list(out["z"]) == [12.0, 11.0, 10.0]
    for i, z in enumerate(out["z"]):
        single = fb.conv_bubbles(out["density"][i], {"BoxSize": L, "HIIdim": D, "Z": float(z)}, LE=False,
                                 quiet=True).cpu().numpy()
        got = out["ionized"][i].cpu().numpy()
        print(f"planck z={z}: x_HI {out['x_HI'][i]:.4f}", flush=True)
        assert np.mean((got == 1) != (single == 1)) <= 1e-5
        both = (got < 1) & (single < 1)
        assert np.allclose(got[both], single[both], rtol=0, atol=1e-5)
    assert out["x_HI"][0] > out["x_HI"][1] > out["x_HI"][2]
    assert 0.0 < out["x_HI"][2] and out["x_HI"][0] < 1.0
