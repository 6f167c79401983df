"""The four stages through their reference entry points and files.

initialize.run() -> perturb_field.run(z) -> find_HII_bubbles.run(opts, [file])
-> delta_T.run(xHfile, IO_DIR), with the reference's defaults (Planck13
spectrum, growth factors and barrier, NGP mass move, HII_FILTER 1, centre-pixel
update) and file names (initialize.py:82,93,109; perturb_field.py:50,108,126;
find_HII_bubbles.py:227-230; delta_T.py:77-78), in a temporary IO_DIR.  Each
saved box is checked against the oracle run on the saved input of its stage."""
import glob
import os
import types

import numpy as np
import pytest
from scipy.special import erfc

from conftest import pkg
from oracle import bubbles, deltat, grf, perturb, sigma

pytestmark = pytest.mark.gpu

DIM, HII, L, Z = 64, 32, 64.0, 10.0


def test_stage_chain_through_files(gpu, tmp_path):
    init = pkg("Programs.initialize")
    pf = pkg("Programs.perturb_field")
    fb = pkg("Programs.find_HII_bubbles")
    dt = pkg("Programs.delta_T")
    pl = pkg("cosmo_files.planck")
    P = pkg("Parameter_files.INIT_PARAMS")
    io = str(tmp_path)
    boxes = os.path.join(io, "Boxes")

    # initialize: the Planck13 table spectrum, reference file names
    out = init.run(DIM=DIM, HII_DIM=HII, BOX_LEN=L, seed=2, out_dir=io)
    names = [f"deltax_z0.00_{DIM}_{L:.0f}Mpc.npy", f"smoothed_deltax_z0.00_{HII}_{L:.0f}Mpc.npy"] + \
            [f"v{m}overddot_{HII}_{L:.0f}Mpc.npy" for m in "xyz"]
    for n in names:
        assert os.path.exists(os.path.join(boxes, n)), n
    want = grf.initialize(DIM, HII, L, 2, pk_table=out["pk_table"])
    dx = np.load(os.path.join(boxes, names[0]))
    assert np.abs(dx - want["deltax"]).max() < 1e-5 * np.abs(want["deltax"]).max()
    print("initialize ok", flush=True)

    # perturb_field.run(z): Planck growth factors, NGP move, velocity boxes at z
    pf.run(Z, DIM=DIM, HII_DIM=HII, BOX_LEN=L, io_dir=io)
    upd_name = os.path.join(boxes, f"updated_smoothed_deltax_z{Z:.2f}_{HII}_{L:.0f}Mpc.npy")
    upd = np.load(upd_name)
    v = [np.load(os.path.join(boxes, n)) for n in names[2:]]
    g, g0 = pl.fgrowth(Z), pl.fgrowth(P.INITIAL_REDSHIFT)
    want_upd = perturb.move_mass_ngp(dx, *v, np.float32(g), np.float32(g0), L)
    assert np.allclose(upd, want_upd, rtol=1e-5, atol=2e-6)
    for m in "xyz":
        assert os.path.exists(os.path.join(boxes, f"updated_v{m}overddot_{HII}_{L:.0f}Mpc.npy"))
    print("perturb_field ok", flush=True)

    # find_HII_bubbles.run: 'updated' input = nonlinear, Planck barrier at Z
    opts = types.SimpleNamespace(DIR=boxes, FILTER_TYPE=1, UPDATE_TYPE=0, Z=Z, QUIET=True)
    ion = fb.run(opts, [upd_name])
    xh_files = glob.glob(os.path.join(boxes, f"xH_nohalos_z{Z:.2f}_nf*_HIIfilter1_*_{HII}_{L:.0f}Mpc.npy"))
    assert len(xh_files) == 1
    xH = np.load(xh_files[0])
    assert np.array_equal(xH, 1 - ion)
    b = pl.barrier(Z)
    radii, _ = sigma.radius_schedule(L, HII)
    den = sigma.denominators(radii, b["sigma_min2"], b["sig0"])
    fc_mean = np.float32(erfc(b["deltac"] / np.sqrt(2 * b["sigma_min2"])))
    gi = np.float32(b["deltac"] / 1.686)
    want_ion = bubbles.find_bubbles_f32(upd, L, radii, den, P.ZETA, b["deltac"], LE=False, growth_inv=gi,
                                        fc_mean=fc_mean)
    _, margin = bubbles.find_bubbles_f64(upd, L, radii, den, P.ZETA, b["deltac"], LE=False,
                                         growth_inv=float(b["deltac"]) / 1.686, fc_mean=float(fc_mean))
    ok = margin > 1e-5
    print(f"find_HII_bubbles: x_HI {xH.mean():.4f}, oracle {1 - want_ion.mean():.4f}", flush=True)
    assert 0.0 < np.mean(ion == 1) < 1.0
    assert np.array_equal((ion == 1)[ok], (want_ion == 1)[ok])
    part = ok & (ion < 1) & (want_ion < 1)
    assert np.allclose(ion[part], want_ion[part], rtol=1e-4, atol=1e-7)

    # delta_T.run: brightness temperature box and P21 against the oracle
    K, k_ave, ps = dt.run(xh_files[0], IO_DIR=io, nbins=20)
    dT_files = glob.glob(os.path.join(boxes, f"delta_T_no_halos_z{Z:.2f}_*_{HII}_{L:.0f}Mpc.npy"))
    assert len(dT_files) == 1
    dT = np.load(dT_files[0])
    want_dT = deltat.delta_T_f32(xH, upd, Z)
    assert np.allclose(dT, want_dT, rtol=1e-6, atol=1e-6)
    edges = deltat.log_edges(L, HII, 20)
    _, want_ps, cnt, _ = deltat.power_spectrum(want_dT.astype(np.float64), L, edges)
    live = cnt > 0
    assert np.allclose(ps[live], want_ps[live], rtol=1e-4, atol=1e-12)


def test_linear_evolution_and_raw_box_input(gpu, tmp_path):
    """perturb_field.run(z) with EVOLVE_DENSITY_LINEARLY: the smoothed z=0
    box times the Planck13 D(z) (perturb_field.py:39-50, 133-138), written
    under the reference's name; find_HII_bubbles.run on the same box stored
    as a raw 21cmFAST float32 file (boxio) gives the result of the .npy
    input."""
    init = pkg("Programs.initialize")
    pf = pkg("Programs.perturb_field")
    fb = pkg("Programs.find_HII_bubbles")
    pl = pkg("cosmo_files.planck")
    boxio = pkg("tocmfastpy.boxio")
    io = str(tmp_path)
    boxes = os.path.join(io, "Boxes")
    init.run(DIM=DIM, HII_DIM=HII, BOX_LEN=L, seed=3, out_dir=io, with_velocity=False)
    upd = pf.run(Z, DIM=DIM, HII_DIM=HII, BOX_LEN=L, io_dir=io, linear=True).cpu().numpy()
    sm = np.load(os.path.join(boxes, f"smoothed_deltax_z0.00_{HII}_{L:.0f}Mpc.npy"))
    assert np.array_equal(upd, perturb.evolve_linear(sm, pl.fgrowth(Z)))
    name = os.path.join(boxes, f"updated_smoothed_deltax_z{Z:.2f}_{HII}_{L:.0f}Mpc")
    assert np.array_equal(np.load(name + ".npy"), upd)
    # the raw file of the same box (no extension): the reader path of find_HII_bubbles.run
    boxio.save_box(name, upd)
    opts = types.SimpleNamespace(DIR=boxes, FILTER_TYPE=1, UPDATE_TYPE=0, Z=Z, QUIET=True)
    a = fb.run(opts, [name + ".npy"])
    b = fb.run(opts, [name])
    assert np.array_equal(a, b) and 0.0 < np.mean(a == 1) < 1.0
