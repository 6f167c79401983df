"""CPU tests of the host side: the C ABI library (loads, exports every symbol
include/cm21.h declares, struct layout, argument checks that return before any
GPU call) and the host logic of the Programs/ modules against the oracle."""
import ctypes
import os
import re
import subprocess

import numpy as np
import pytest

from conftest import ROOT, pkg

HEADER = os.path.join(ROOT, "include", "cm21.h")


def header_functions():
    """Names of the functions include/cm21.h declares."""
    text = re.sub(r"/\*.*?\*/", "", open(HEADER).read(), flags=re.S)
    return sorted(set(re.findall(r"^\s*(?:const\s+)?[a-z_0-9]+\s*\**\s*(cm21_[A-Za-z0-9_]+)\s*\(", text, flags=re.M)))


@pytest.fixture(scope="module")
def cdll():
    lib_mod = pkg("_lib")
    if not os.path.exists(lib_mod.LIB_PATH):
        spec = __import__("importlib.util").util.spec_from_file_location(
            "cm21_build", os.path.join(ROOT, "21cmpy_amd", "build.py"))
        mod = __import__("importlib.util").util.module_from_spec(spec)
        spec.loader.exec_module(mod)
        mod.build()
    return lib_mod.load()


def test_header_parse_finds_all_entry_points():
    names = header_functions()
    assert "cm21_find_bubbles" in names and "cm21_power_spectrum" in names
    assert "cm21_delta_T" in names and len(names) == len(pkg("_lib").SIGNATURES) >= 45


def test_library_exports_every_declared_symbol(cdll):
    missing = [n for n in header_functions() if not hasattr(cdll, n)]
    assert not missing, missing


def test_bindings_cover_header_exactly():
    lib_mod = pkg("_lib")
    assert sorted(lib_mod.SIGNATURES) == header_functions()


def test_nm_dynamic_exports(cdll):
    lib_mod = pkg("_lib")
    out = subprocess.run(["nm", "-D", "--defined-only", lib_mod.LIB_PATH], capture_output=True, text=True,
                         check=True).stdout
    exported = {line.split()[-1] for line in out.splitlines() if " T " in line}
    assert set(header_functions()) <= exported


@pytest.mark.parametrize("cname,pyname", [("cm21_bubble_params", "BubbleParams"), ("cm21_batch_box", "BatchBox"),
                                           ("cm21_batch_group", "BatchGroup"), ("cm21_unet_conv", "UnetConv"),
                                           ("cm21_unet_wgrad", "UnetWgrad")])
def test_struct_layout_matches_c(tmp_path, cname, pyname):
    """Each ctypes structure has the C struct's size and field offsets (gcc on the header)."""
    lib_mod = pkg("_lib")
    cls = getattr(lib_mod, pyname)
    fields = [f for f, _ in cls._fields_]
    src = tmp_path / "layout.c"
    body = "".join(f'    printf("{f} %zu\\n", offsetof({cname}, {f}));\n' for f in fields)
    src.write_text('#include <stdio.h>\n#include <stddef.h>\n#include "cm21.h"\nint main(void){\n'
                   f'    printf("sizeof %zu\\n", sizeof({cname}));\n' + body + "    return 0;\n}\n")
    exe = tmp_path / "layout"
    subprocess.run(["gcc", "-std=c99", "-I", os.path.join(ROOT, "include"), str(src), "-o", str(exe)], check=True)
    got = dict(line.split() for line in subprocess.run([str(exe)], capture_output=True, text=True,
                                                        check=True).stdout.splitlines())
    assert int(got["sizeof"]) == ctypes.sizeof(cls)
    for f in fields:
        assert int(got[f]) == getattr(cls, f).offset, f


def test_abi_version_and_argument_errors(cdll):
    assert cdll.cm21_abi_version() == 1
    h = ctypes.c_void_p()
    assert cdll.cm21_plan_create(None, 64, 1) == -1
    assert b"NULL" in cdll.cm21_last_error()
    assert cdll.cm21_plan_create(ctypes.byref(h), 63, 1) == -1
    assert b"D=63" in cdll.cm21_last_error()
    assert cdll.cm21_plan_create(ctypes.byref(h), 64, 0) == -1
    assert h.value is None


def test_workspace_sizes(cdll):
    """Workspace queries are pure host arithmetic: they grow with D and with the
    number of radii; without the sphere paint they cover the spectrum and the
    five intermediates of the pipelined loop (rows padded to 16 modes) and the
    mask; with it, the spectrum, the filtered spectrum, the smoothed box and
    the two uint32 distance boxes of the paint."""
    D = 256
    nc = D * D * (D // 2 + 1)
    nt = D * D * ((D // 2 + 1 + 15) // 16 * 16)
    w40 = cdll.cm21_find_bubbles_workspace_bytes(D, 40, 0)
    assert w40 >= 6 * 8 * nt + D ** 3 // 8
    assert cdll.cm21_find_bubbles_workspace_bytes(D, 41, 0) >= w40
    assert cdll.cm21_find_bubbles_workspace_bytes(128, 40, 0) < w40
    assert cdll.cm21_find_bubbles_workspace_bytes(D, 40, 1) >= 8 * nc + 8 * nt + 4 * D ** 3 + 2 * 4 * D ** 3
    assert cdll.cm21_reduce_workspace_bytes() >= 8
    assert cdll.cm21_pk_workspace_bytes(D, 40) > 0


@pytest.mark.parametrize("L,D,py2", [(100.0, 64, False), (300.0, 256, False), (500.0, 512, False),
                                      (150.0, 96, False), (200.0, 128, False), (250.0, 160, False),
                                      (300.0, 256, True), (75.0, 64, True)])
def test_radius_schedule_matches_oracle(L, D, py2):
    fb = pkg("Programs.find_HII_bubbles")
    from oracle import sigma
    r, cell = fb.radius_schedule(L, D, py2_int_scale=py2)
    ro, cello = sigma.radius_schedule(L, D, py2_int_scale=py2)
    assert r.dtype == np.float32
    np.testing.assert_array_equal(r, ro)
    assert cell == cello


def test_barrier_denominators_match_oracle_and_raise():
    fb = pkg("Programs.find_HII_bubbles")
    from oracle import sigma
    r, _ = fb.radius_schedule(300.0, 256)
    fn = lambda R: sigma.sig0_discrete(R, 256, 300.0)  # noqa: E731
    np.testing.assert_array_equal(fb.barrier_denominators(r, 4.0, fn), sigma.denominators(r, 4.0, fn))
    with pytest.raises(ValueError):
        fb.barrier_denominators(r, 0.3, fn)  # S(R) of the smallest radius exceeds sigma_min^2


def test_default_sig0_matches_oracle():
    fb = pkg("Programs.find_HII_bubbles")
    from oracle import sigma
    fn = fb._default_sig0(64, 100.0)
    for R in (30.0, 10.0, 3.0, 1.6):
        assert fn(R) == pytest.approx(sigma.sig0_discrete(R, 64, 100.0), rel=1e-12)


def test_sig0_memo_returns_the_computed_values():
    """S(R) is kept per radius on each spectrum object (the radius loop asks
    for the same radii on every call); the kept values are the computed ones,
    and the default spectrum of a box shape is built once."""
    fb = pkg("Programs.find_HII_bubbles")
    sig = pkg("cosmo_files.sigmas")
    pl = pkg("cosmo_files.planck")
    r, _ = fb.radius_schedule(300.0, 256)
    spec = sig.PowerLawSpectrum(256, 300.0)
    first = [spec.sig0_discrete(float(R)) for R in r]
    again = [spec.sig0_discrete(np.float32(R)) for R in r]
    fresh = [sig.PowerLawSpectrum(256, 300.0).sig0_discrete(float(R)) for R in r[:5]]
    assert first == again and first[:5] == fresh
    assert fb._default_sig0(256, 300).__self__ is fb._default_sig0(256.0, 300.0).__self__
    tab = pl.planck_spectrum()
    assert tab.sig0(5.0) == tab.sig0(5.0) == pl.sig0_trapz(5.0, tab.del2)


def test_radius_schedule_uses_callers_scale():
    """conv_bubbles' scale argument sets the cell size of the stop rule
    (find_HII_bubbles.py:52-54,123): the radius count follows the oracle's
    schedule with the same scale, not BoxSize/HIIdim."""
    from oracle import sigma
    fb = pkg("Programs.find_HII_bubbles")
    for L, D, scale in ((300.0, 256, 3.0), (100.0, 64, 0.5), (200.0, 128, 2.4)):
        r, cell = fb.radius_schedule(L, D, cell=scale)
        want, wcell = sigma.radius_schedule(L, D, scale=scale)
        assert cell == wcell == scale and np.array_equal(r, want)
        default, _ = fb.radius_schedule(L, D)
        assert len(r) != len(default)
        radii, _, _ = fb.bubble_setup(D, L, scale, 1.686, 50.0, lambda R: 0.1)
        assert np.array_equal(radii, want)


def test_conv_bubbles_fails_loudly_without_gpu():
    import torch
    if torch.cuda.is_available():
        pytest.skip("a GPU is visible")
    fb = pkg("Programs.find_HII_bubbles")
    with pytest.raises(RuntimeError, match="GPU"):
        fb.conv_bubbles(np.zeros((16, 16, 16), np.float32), {"BoxSize": 50, "HIIdim": 16}, LE=True,
                        deltac=1.686, sigma_min2=0.5)


def test_parse_filename_reference_names():
    boxio = pkg("tocmfastpy.boxio")
    d = boxio.parse_filename("Boxes/updated_smoothed_deltax_z008.00_256_300Mpc")
    assert d["type"] == "density" and d["z"] == 8.0 and d["HIIdim"] == 256 and d["BoxSize"] == 300
    d = boxio.parse_filename("xH_nohalos_z010.00_nf0.512345_eff30.0_HIIfilter1_Mmin3.4e+08_RHIImax30_256_300Mpc")
    assert d["type"] == "xh" and d["nf"] == pytest.approx(0.512345) and d["HIIfilter"] == 1
    assert d["eff"] == 30.0 and d["RHIIfilter"] == 30.0 and d["Mmin"] == pytest.approx(3.4e8)
    d = boxio.parse_filename("deltax_z0.00_512_500Mpc")
    assert d["type"] == "init_x" and d["HIIdim"] == 512 and d["BoxSize"] == 500
    assert boxio.parse_filename("Boxes/")["type"] is False


def test_open_box_padded_and_plain(tmp_path):
    boxio = pkg("tocmfastpy.boxio")
    D = 8
    rng = np.random.default_rng(0)
    box = rng.standard_normal((D, D, D)).astype(np.float32)
    plain = tmp_path / "deltax_z0.00_8_10Mpc"
    box.tofile(plain)
    np.testing.assert_array_equal(boxio.open_box(str(plain), D), box)
    padded = np.zeros((D, D, 2 * (D // 2 + 1)), np.float32)
    padded[:, :, :D] = box
    pfile = tmp_path / "deltak_z0.00_8_10Mpc"
    padded.tofile(pfile)
    np.testing.assert_array_equal(boxio.trim_box(boxio.open_box(str(pfile), D)), box)


def test_output_names():
    fb = pkg("Programs.find_HII_bubbles")
    dt = pkg("Programs.delta_T")
    name = fb.xH_filename(8.0, 0.25, 1, 256, 300)
    assert name == "xH_nohalos_z8.00_nf0.250000_eff40.0_HIIfilter1_Tvirmin10000.0_RHIImax30_256_300Mpc.npy"
    assert boxio_roundtrip(name)["nf"] == 0.25
    n2 = dt.delta_T_name(8.0, 0.25, 30.0, 12.5, 256, 300)
    assert n2.startswith("delta_T_no_halos_z8.00_nf0.250000_useTs0") and n2.endswith("_aveTb12.50_256_300Mpc.npy")


def boxio_roundtrip(name):
    return pkg("tocmfastpy.boxio").parse_filename(name)


def test_bin_lut_matches_oracle_binning():
    gpu = pkg("gpu")
    dt = pkg("Programs.delta_T")
    from oracle import deltat
    D, L = 32, 100.0
    edges = dt.log_edges(L, D, 40)
    np.testing.assert_allclose(edges, deltat.log_edges(L, D, 40), rtol=0, atol=0)
    lut = gpu.bin_lut(D, L, edges)
    m = np.arange(3 * (D // 2) ** 2 + 1)
    ref = deltat.bin_of_nsq(m, L, edges)
    assert lut[0] == -1
    np.testing.assert_array_equal(lut[1:], ref[1:])


def test_tb_constant_matches_oracle():
    dt = pkg("Programs.delta_T")
    from oracle import deltat
    for z in (6.0, 8.0, 12.0):
        assert float(dt.tb_constant(z)) == pytest.approx(deltat.tb_constant(z), rel=1e-6)


def test_init_params_values():
    P = pkg("Parameter_files.INIT_PARAMS")
    assert (P.R_BUBBLE_MAX, P.R_BUBBLE_MIN, P.DELTA_R_FACTOR) == (30.0, 0.5, 1.1)
    assert P.ZETA == 40.0 and P.HII_FILTER == 1 and P.FIND_BUBBLE_ALGORITHM == 0
    assert P.L_FACTOR == np.float32(0.620350491) and P.ION_Tvir_MIN == 1.0e4


def test_batched_workspace_query(cdll):
    """The batched size query is host arithmetic: it accepts the 64-box
    manifest's groups and rejects unsupported shapes with 0."""
    import json
    lib_mod = pkg("_lib")
    fb = pkg("Programs.find_HII_bubbles")
    man = json.load(open(os.path.join(ROOT, "21cmpy_amd", "Parameter_files", "batch64_manifest.json")))
    groups = {}
    for b in man["boxes"]:
        groups.setdefault(b["D"], []).append(b)
    assert sorted(groups) == [96, 128, 160]
    arr = (lib_mod.BatchGroup * 3)()
    keep = []
    for k, D in enumerate(sorted(groups)):
        radii, _ = fb.radius_schedule(1.5625 * D, D)
        assert len(radii) == 37
        keep.append(radii)
        arr[k] = lib_mod.BatchGroup(D=D, nbox=len(groups[D]), box_len=1.5625 * D, n_radii=len(radii),
                                    radii=radii.ctypes.data_as(lib_mod.c_float_p), tile_columns=-1, tile_lines=-1)
    prm = lib_mod.BubbleParams(filter_type=1, ref_fallthrough=1, linear_input=1, update=0, fused=1)
    n = cdll.cm21_find_bubbles_batched_workspace_bytes(arr, 3, ctypes.byref(prm))
    spectra = sum(len(groups[D]) * 8 * D * D * (D // 2 + 1) for D in groups)
    assert spectra < n < 40 * spectra
    prm.filter_type = 0
    assert cdll.cm21_find_bubbles_batched_workspace_bytes(arr, 3, ctypes.byref(prm)) == 0
    prm.filter_type = 1
    arr[0].D = 100
    assert cdll.cm21_find_bubbles_batched_workspace_bytes(arr, 3, ctypes.byref(prm)) == 0
    assert cdll.cm21_batch_tile_count(128, 0) >= 1 and cdll.cm21_batch_tile_count(100, 0) == -1
    assert cdll.cm21_batch_tile_name(128, 1, 0) is not None


def test_batch_manifest_is_reproducible():
    """The committed 64-box manifest is what its fixed generator draws."""
    import json
    bm = pkg("Programs.find_HII_bubbles_batched")
    assert json.load(open(bm.MANIFEST)) == json.loads(json.dumps(bm.make_manifest()))
    man = bm.load_manifest()
    for b in man["boxes"]:
        assert b["box_len"] == 1.5625 * b["D"] and 10 <= b["zeta"] <= 60 and 2 <= b["sigma_min2"] <= 6
        assert b["z"] in (7.0, 8.0, 9.0, 10.0)
    assert [b["seed"] for b in man["boxes"]] == list(range(64))


def test_batch_box_parameters():
    bm = pkg("Programs.find_HII_bubbles_batched")
    from scipy.special import erfc
    deltac, ginv, fc = bm.box_parameters({"z": 8.0, "sigma_min2": 4.0})
    assert deltac == pytest.approx(1.686 / 0.12)
    assert float(ginv) == pytest.approx(1 / 0.12, rel=1e-6)
    assert float(fc) == pytest.approx(erfc(deltac / np.sqrt(8.0)), rel=1e-6)
    assert bm.box_parameters({"deltac": 2.0, "sigma_min2": 1.0})[0] == 2.0
    assert bm.linear_growth(8.0) == pytest.approx(0.12)


def test_batch_tiles_file_names_valid_candidates(cdll):
    """The committed autotuned tiles name existing candidates of each D."""
    bm = pkg("Programs.find_HII_bubbles_batched")
    tiles = bm.load_tiles()
    assert tiles
    for key, v in tiles.items():
        D = int(key.split("_")[0][1:])
        assert 0 <= v["columns"] < cdll.cm21_batch_tile_count(D, 0)
        assert 0 <= v["lines"] < cdll.cm21_batch_tile_count(D, 1)
        assert v["chosen"][0] == cdll.cm21_batch_tile_name(D, 0, v["columns"]).decode()


def test_surrogate_specs_and_weights_file():
    sur = pkg("surrogate")
    a = sur.surrogate_specs(range(1000, 1010))
    assert a == sur.surrogate_specs(range(1000, 1010))
    for s in a:
        assert 10 <= s["zeta"] <= 60 and 2 <= s["sigma_min2"] <= 6 and s["deltac"] == 4.0 and s["D"] == 128
    params, meta = sur.UNet3D.load_file(sur.WEIGHTS)
    shapes = dict(sur.param_shapes())
    assert set(params) == set(shapes)
    for k, v in params.items():
        assert v.shape == shapes[k] and np.all(np.isfinite(v))
    assert meta["D"] == 128 and meta["deltac"] == 4.0
