"""GPU parity of the batched find_HII_bubbles (cm21_find_bubbles_batched)
against the FP32 oracle and against the single-box loop.

Bar: ionized flags identical except cells whose threshold margin
|zeta f_coll - 1| falls below 1e-5 at some radius (FP64 oracle); partially
ionized values within 1e-4 relative.  Against the single-box loop, which
differs only in the forward transform (a batched rocFFT plan), at most 1e-5 of
the cells may change flag and values agree within 1e-5."""
import numpy as np
import pytest
import torch

from conftest import pkg
from oracle import sigma
from test_gpu_bubbles import _compare, _oracle

pytestmark = pytest.mark.gpu


def _box(D, L, seed, scale=1.0):
    """Synthetic GRF density box on the device (product GRF path)."""
    g = pkg("gpu")
    sp = pkg("cosmo_files.sigmas").PowerLawSpectrum(D, L)
    d = g.c2r(g.grf_init(D, L, sp.amp, seed))
    return (d * scale).contiguous() if scale != 1.0 else d


def _specs_linear():
    # non-degenerate linear input: barrier 1.686, sigma_min^2 a little above S(R)
    return [
        {"D": 64, "box_len": 100.0, "zeta": 30.0, "deltac": 1.686, "sigma_min2": 0.5, "seed": 11},
        {"D": 96, "box_len": 150.0, "zeta": 18.0, "deltac": 1.686, "sigma_min2": 0.62, "seed": 12},
        {"D": 64, "box_len": 100.0, "zeta": 55.0, "deltac": 1.75, "sigma_min2": 0.8, "seed": 13},
        {"D": 96, "box_len": 150.0, "zeta": 40.0, "deltac": 1.6, "sigma_min2": 0.45, "seed": 14},
        {"D": 64, "box_len": 100.0, "zeta": 12.0, "deltac": 1.686, "sigma_min2": 0.55, "seed": 15},
    ]


def _with_sig0(specs):
    out = []
    for s in specs:
        s = dict(s)
        D, L = s["D"], s["box_len"]
        s["sig0"] = (lambda R, D=D, L=L: sigma.sig0_discrete(R, D, L))
        out.append(s)
    return out


def _single(delta, s, LE, fil=1):
    fb = pkg("Programs.find_HII_bubbles")
    return fb.conv_bubbles(delta, {"BoxSize": s["box_len"], "HIIdim": s["D"]}, fil=fil, LE=LE, quiet=True,
                           deltac=s["deltac"], sigma_min2=s["sigma_min2"], zeta=s["zeta"], sig0=s["sig0"],
                           fused=True).cpu().numpy()


def _close_to_single(got, want):
    flips = np.mean((got == 1) != (want == 1))
    assert flips <= 1e-5, f"flag flips {flips}"
    both = (got < 1) & (want < 1)
    assert np.allclose(got[both], want[both], rtol=0, atol=1e-5)


@pytest.mark.parametrize("fil", [1, 2])
def test_batched_linear_matches_oracle_and_single(gpu, fil):
    bm = pkg("Programs.find_HII_bubbles_batched")
    specs = _with_sig0(_specs_linear())
    deltas = [_box(s["D"], s["box_len"], s["seed"]) for s in specs]
    bb = bm.BatchedBubbles(specs, fil=fil, LE=True, graph=False, tiles=None)
    assert [g[0] for g in bb.groups] == [64, 96]
    bb.set_deltas(deltas)
    bb.run()
    torch.cuda.synchronize()
    for i, (s, d) in enumerate(zip(specs, deltas)):
        got = bb.ionized(i).cpu().numpy()
        print(f"box {i} D={s['D']} zeta={s['zeta']} x_ion={np.mean(got == 1):.3f}", flush=True)
        _close_to_single(got, _single(d, s, True, fil))
        if i < 2:
            want, margin = _oracle(d.cpu().numpy(), s["box_len"], s["sigma_min2"], s["zeta"], True, s["deltac"],
                                   fil)
            assert 0.02 < np.mean(want == 1) < 0.98
            _compare(got, want, margin)
    xhi = bb.x_HI()
    for i in range(len(specs)):
        assert xhi[i] == pytest.approx(1.0 - float(bb.ionized(i).double().mean()), abs=1e-9)


def test_batched_nonlinear_matches_single(gpu):
    bm = pkg("Programs.find_HII_bubbles_batched")
    growth = 0.12
    # zeta * erfc(deltac / sqrt(2 sigma_min^2)) near 1: partly ionized boxes
    specs = _with_sig0([
        {"D": 64, "box_len": 100.0, "zeta": 30.0, "deltac": 1.686 / growth, "sigma_min2": 44.0, "seed": 21},
        {"D": 64, "box_len": 100.0, "zeta": 38.0, "deltac": 1.686 / growth, "sigma_min2": 40.0, "seed": 22},
        {"D": 128, "box_len": 200.0, "zeta": 22.0, "deltac": 1.686 / growth, "sigma_min2": 50.0, "seed": 23},
    ])
    deltas = [_box(s["D"], s["box_len"], s["seed"], growth) for s in specs]
    bb = bm.BatchedBubbles(specs, fil=1, LE=False, graph=False, tiles=None)
    bb.set_deltas(deltas)
    bb.run()
    torch.cuda.synchronize()
    for i, (s, d) in enumerate(zip(specs, deltas)):
        got = bb.ionized(i).cpu().numpy()
        print(f"box {i} D={s['D']} x_ion={np.mean(got == 1):.4f}", flush=True)
        assert 0.001 < np.mean(got == 1) < 0.999
        _close_to_single(got, _single(d, s, False))
    want, margin = _oracle(deltas[0].cpu().numpy(), 100.0, 44.0, 30.0, False, 1.686 / growth)
    _compare(bb.ionized(0).cpu().numpy(), want, margin)


def test_batched_every_tile_agrees(gpu):
    """Every columns/lines candidate gives the default's flags (the DFT
    factorisation changes rounding only) for D in {96, 128, 160}."""
    bm = pkg("Programs.find_HII_bubbles_batched")
    lib = pkg("_lib").load()
    specs = _with_sig0([
        {"D": 96, "box_len": 150.0, "zeta": 25.0, "deltac": 1.686, "sigma_min2": 0.5, "seed": 31},
        {"D": 128, "box_len": 200.0, "zeta": 35.0, "deltac": 1.686, "sigma_min2": 0.6, "seed": 32},
        {"D": 160, "box_len": 250.0, "zeta": 20.0, "deltac": 1.686, "sigma_min2": 0.5, "seed": 33},
    ])
    bb = bm.BatchedBubbles(specs, fil=1, LE=True, graph=False, tiles=None)
    bb.set_deltas([_box(s["D"], s["box_len"], s["seed"]) for s in specs])
    bb.run()
    ref = [bb.ionized(i).cpu().numpy() for i in range(3)]
    for i in range(3):
        assert 0.02 < np.mean(ref[i] == 1) < 0.98
    for k, (D, L, idx) in enumerate(bb.groups):
        key = list(bb.tiles())[k]
        for c in range(lib.cm21_batch_tile_count(D, 0)):
            for l in range(lib.cm21_batch_tile_count(D, 1)):
                print(f"D={D} columns {c} lines {l}", flush=True)
                bb.set_tiles({key: (c, l)})
                bb.run()
                _close_to_single(bb.ionized(idx[0]).cpu().numpy(), ref[idx[0]])
        bb.set_tiles({key: (-1, -1)})


def test_batched_graph_replay_and_autotune(gpu, tmp_path):
    bm = pkg("Programs.find_HII_bubbles_batched")
    specs = _with_sig0(_specs_linear()[:4])
    deltas = [_box(s["D"], s["box_len"], s["seed"]) for s in specs]
    eager = bm.BatchedBubbles(specs, graph=False, tiles=None)
    eager.set_deltas(deltas)
    eager.run()
    want = [eager.ionized(i).clone() for i in range(4)]
    torch.cuda.synchronize()
    st = torch.cuda.Stream()
    with torch.cuda.stream(st):
        bb = bm.BatchedBubbles(specs, graph=True, tiles=None, stream=st)
        bb.set_deltas(deltas)
        res = bb.autotune(reps=2, save_path=str(tmp_path / "tiles.json"))
        assert len(res) == 2 and all(v["columns"] >= 0 and v["lines"] >= 0 for v in res.values())
        assert all(len(v["ms"]) == len(v["names"]) and min(v["ms"]) > 0 for v in res.values())
        bb.set_deltas(deltas)  # the tuner leaves the outputs unspecified
        for rep in range(3):
            bb.run()
            st.synchronize()
            for i in range(4):
                got = bb.ionized(i).cpu().numpy()
                _close_to_single(got, want[i].cpu().numpy())
            # eager work between replays
            torch.cuda.synchronize()
    saved = bm.load_tiles(str(tmp_path / "tiles.json"))
    assert set(saved) == set(res)


def test_batched_fully_ionized_early_exit(gpu):
    """Literal parameters of the 64^3 configuration (sigma_min^2 = 4, zeta = 30,
    deltac = 1.686): every cell ionizes at the first radius; the persistent
    pass stops early and paints 1."""
    bm = pkg("Programs.find_HII_bubbles_batched")
    specs = _with_sig0([{"D": 64, "box_len": 100.0, "zeta": 30.0, "deltac": 1.686, "sigma_min2": 4.0, "seed": 0},
                        {"D": 64, "box_len": 100.0, "zeta": 30.0, "deltac": 1.686, "sigma_min2": 0.5, "seed": 1}])
    bb = bm.BatchedBubbles(specs, graph=False, tiles=None)
    bb.set_deltas([_box(64, 100.0, s["seed"]) for s in specs])
    bb.run()
    assert torch.all(bb.ionized(0) == 1)
    got = bb.ionized(1).cpu().numpy()
    assert 0.02 < np.mean(got == 1) < 0.98


def test_batched_manifest_shapes(gpu):
    """The 64-box manifest builds (groups by D, workspace fits) and runs one
    pass; its literal parameters leave every box neutral (barrier 1.686/D(z)
    against per-cell sigma 0.6), so x_HI = 1 for all boxes."""
    bm = pkg("Programs.find_HII_bubbles_batched")
    man = bm.load_manifest()
    specs = man["boxes"]
    assert len(specs) == 64 and {s["D"] for s in specs} == {96, 128, 160}
    bb = bm.BatchedBubbles(specs, graph=True, stream=torch.cuda.Stream())
    for i, s in enumerate(specs):
        bb.set_delta(i, _box(s["D"], s["box_len"], s["seed"]))
    bb.run()
    torch.cuda.synchronize()
    assert np.allclose(bb.x_HI(), 1.0)


def test_batched_errors(gpu):
    bm = pkg("Programs.find_HII_bubbles_batched")
    with pytest.raises(ValueError):
        bm.BatchedBubbles([{"D": 64, "box_len": 100.0, "zeta": 30.0, "deltac": 1.686, "sigma_min2": 0.1}])
    with pytest.raises(ValueError):
        bm.BatchedBubbles([{"D": 64, "box_len": 100.0, "zeta": 30.0, "deltac": 1.686, "sigma_min2": 0.5}], fil=0)
    lib = pkg("_lib")
    with pytest.raises(lib.CM21Error):
        bb = bm.BatchedBubbles([{"D": 64, "box_len": 100.0, "zeta": 30.0, "deltac": 1.686, "sigma_min2": 0.5}],
                               graph=False, tiles=None)
        bb._garr[0].tile_lines = 99
        bb.run()


def test_batched_groups_sharing_a_plan(gpu):
    """Two groups with the same D and box count but different box_len share
    one rocFFT plan (and its work buffer); they must not run their forward
    transforms at the same time.  Each box equals the single-box loop, and a
    run on a stream of its own, with the boxes loaded on the caller's stream,
    gives the same result."""
    bm = pkg("Programs.find_HII_bubbles_batched")
    specs = _with_sig0([
        {"D": 64, "box_len": 100.0, "zeta": 30.0, "deltac": 1.686, "sigma_min2": 0.5, "seed": 21},
        {"D": 64, "box_len": 80.0, "zeta": 25.0, "deltac": 1.686, "sigma_min2": 0.6, "seed": 22},
        {"D": 64, "box_len": 100.0, "zeta": 40.0, "deltac": 1.7, "sigma_min2": 0.55, "seed": 23},
        {"D": 64, "box_len": 80.0, "zeta": 20.0, "deltac": 1.65, "sigma_min2": 0.65, "seed": 24},
    ])
    deltas = [_box(s["D"], s["box_len"], s["seed"]) for s in specs]
    for stream in (None, torch.cuda.Stream()):
        bb = bm.BatchedBubbles(specs, LE=True, graph=stream is not None, tiles=None, stream=stream)
        assert [(g[0], len(g[2])) for g in bb.groups] == [(64, 2), (64, 2)]
        assert bb._groups[0]["plan"] is bb._groups[1]["plan"]
        bb.set_deltas(deltas)
        for _ in range(2):
            bb.run()
            for i, (s, d) in enumerate(zip(specs, deltas)):
                got = bb.ionized(i).cpu().numpy()
                print(f"stream={stream is not None} box {i} L={s['box_len']} x_ion={np.mean(got == 1):.3f}",
                      flush=True)
                assert 0.01 < np.mean(got == 1) < 0.99
                _close_to_single(got, _single(d, s, True))
