"""cm21_find_bubbles_noalloc (the loop with its workspace in the library's
static device array) against cm21_find_bubbles with a caller workspace,
through the C ABI: identical kernels, so identical outputs bit for bit, on
two calls in a row with the second box smaller, for the fused nonlinear,
fused linear and generic sphere-paint paths; boxes too large for the array
are rejected with CM21_ERR_NOMEM."""
import ctypes

import numpy as np
import pytest
import torch

from conftest import pkg

pytestmark = pytest.mark.gpu


def _box(D, L, seed, scale):
    g = pkg("gpu")
    sp = pkg("cosmo_files.sigmas").PowerLawSpectrum(D, L)
    return (g.c2r(g.grf_init(D, L, sp.amp, seed)) * scale).contiguous()


def _setup(D, L, LE, update=0, fused=1):
    lib = pkg("_lib")
    fb = pkg("Programs.find_HII_bubbles")
    sp = pkg("cosmo_files.sigmas").PowerLawSpectrum(D, L)
    growth = 1.0 if LE else 0.5
    deltac = 1.686 / growth
    smin = 0.6 if LE else 1.2
    radii, denoms, fc_mean = fb.bubble_setup(D, L, None, deltac, smin, sp.sig0_discrete)
    prm = lib.BubbleParams(box_len=L, deltac=float(np.float32(deltac)), zeta=30.0,
                           growth_inv=float(np.float32(deltac / 1.686)), fc_mean=float(fc_mean),
                           linear_input=1 if LE else 0, filter_type=1, ref_fallthrough=1, update=update,
                           sphere_periodic=0, sphere_r_scale=1.0, fused=fused, graph=0)
    return prm, np.ascontiguousarray(radii, np.float32), np.ascontiguousarray(denoms, np.float32), growth


def _run_both(D, L, LE, seed, update=0, fused=1):
    lib = pkg("_lib")
    prm, radii, denoms, growth = _setup(D, L, LE, update, fused)
    delta = _box(D, L, seed, growth)
    plan = lib.plan(D, 1)
    st = torch.cuda.current_stream()
    rp, dp = radii.ctypes.data_as(lib.c_float_p), denoms.ctypes.data_as(lib.c_float_p)
    nbytes = lib.query("cm21_find_bubbles_workspace_bytes", D, len(radii), update)
    assert 0 < nbytes <= (64 << 20)
    work = torch.empty(nbytes, dtype=torch.uint8, device="cuda")
    want = torch.empty_like(delta)
    lib.call("cm21_find_bubbles", plan.handle, lib.ptr(delta), lib.ptr(want), rp, dp, len(radii), prm,
             lib.ptr(work), nbytes, lib.stream_handle(st))
    got = torch.full_like(delta, -7.0)
    lib.call("cm21_find_bubbles_noalloc", plan.handle, lib.ptr(delta), lib.ptr(got), rp, dp, len(radii), prm,
             lib.stream_handle(st))
    torch.cuda.synchronize()
    return got.cpu().numpy(), want.cpu().numpy()


@pytest.mark.parametrize("case", ["fused_nonlinear", "fused_linear", "generic_sphere"])
def test_noalloc_equals_workspace_entry_point(gpu, case):
    LE = case != "fused_nonlinear"
    update = 1 if case == "generic_sphere" else 0
    fused = 0 if case == "generic_sphere" else 1
    # two calls in a row, the second on a smaller box (the array keeps the
    # first call's contents: the loop must clear what it needs itself)
    for D, L, seed in ((128, 200.0, 3), (64, 100.0, 4)):
        got, want = _run_both(D, L, LE, seed, update, fused)
        ion = float(np.mean(want == 1))
        print(f"{case} D={D}: ionized fraction {ion:.3f}", flush=True)
        assert ion > 0.0 and (update or ion < 1.0)
        assert np.array_equal(got, want)


def test_noalloc_rejects_boxes_beyond_the_array(gpu):
    lib = pkg("_lib")
    prm, radii, denoms, _ = _setup(256, 300.0, True)
    plan = lib.plan(256, 1)
    nbytes = lib.query("cm21_find_bubbles_workspace_bytes", 256, len(radii), 0)
    assert nbytes > (64 << 20)
    d = torch.zeros((256, 256, 256), device="cuda")
    out = torch.empty_like(d)
    with pytest.raises(lib.CM21Error, match="static scratch") as e:
        lib.call("cm21_find_bubbles_noalloc", plan.handle, lib.ptr(d), lib.ptr(out),
                 radii.ctypes.data_as(lib.c_float_p), denoms.ctypes.data_as(lib.c_float_p), len(radii), prm,
                 lib.stream_handle(torch.cuda.current_stream()))
    assert e.value.status == -5
