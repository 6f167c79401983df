"""GPU parity of the U-Net surrogate's layer kernels and of the whole
forward/backward pass against the CPU oracle (oracle/unet.py, torch float64).

Inputs of each layer test are bf16-exact, so a bf16 output may differ from
the float64 oracle by its own rounding (half a bf16 ulp, 2^-8 relative) plus
fp32 accumulation; fp32 weight gradients agree to 1e-3 of their largest
entry.  The whole network is compared with the oracle that rounds at the same
points as the GPU (bf16=True)."""
import ctypes

import numpy as np
import pytest
import torch

from conftest import pkg
from oracle import unet

pytestmark = pytest.mark.gpu

HALF_ULP = 2.0 ** -8


def _bf(shape, scale=1.0, seed=0):
    g = torch.Generator().manual_seed(seed)
    return (torch.randn(shape, generator=g, dtype=torch.float64) * scale).to(torch.bfloat16).cuda()


def _f(shape, scale=1.0, seed=0):
    g = torch.Generator().manual_seed(seed)
    return (torch.randn(shape, generator=g, dtype=torch.float64) * scale).to(torch.float32).cuda()


def _d(t):
    return t.detach().cpu().to(torch.float64)


def _vp(t):
    return None if t is None else ctypes.c_void_p(t.data_ptr())


def _fp(t):
    return None if t is None else ctypes.cast(ctypes.c_void_p(t.data_ptr()), ctypes.POINTER(ctypes.c_float))


def _close_bf16(got, ref, what):
    got, ref = _d(got), ref.to(torch.float64)
    assert bool(torch.isfinite(got).all()), f"{what}: non-finite values"
    tol = HALF_ULP * ref.abs() + 1e-5 * float(ref.abs().max()) + 1e-30
    bad = ((got - ref).abs() > tol)
    assert not bool(bad.any()), f"{what}: {int(bad.sum())} of {bad.numel()} outside, max err " \
                                f"{float((got - ref).abs().max()):.3g}"


def _close_f32(got, ref, what, rel=1e-3):
    got, ref = _d(got), ref.to(torch.float64)
    err = float((got - ref).abs().max())
    assert err <= rel * float(ref.abs().max()) + 1e-12, f"{what}: max err {err:.3g} vs scale {float(ref.abs().max()):.3g}"


def _conv_call(lib, B, D, cin, cout, x0, w, y0, x1=None, c0=None, mask=None, bias=None, relu=True, y1=None, o0=None,
               per_sample=False, algo=0):
    c = lib.UnetConv(B=B, D=D, cin=cin, cout=cout, c0=c0 or cin, o0=o0 or cout, x0=_vp(x0), x1=_vp(x1),
                     mask=_vp(mask), w=_vp(w), bias=_fp(bias), bias_per_sample=1 if per_sample else 0,
                     relu=1 if relu else 0, y0=_vp(y0), y1=_vp(y1), algo=algo)
    lib.call("cm21_unet_conv3_fwd", ctypes.byref(c), lib.stream_handle())


# D < 64: one 2 x 2 x 16 tile per block; D >= 64: the 4 x 4 x 16 (or
# 4 x 2 x 16) tile per block (algo 1) and the sliding window over 8 (D = 64)
# or 16 planes (D >= 128, algo 2); every D: the tile of 16 input channels at
# a time (algo 3)
SHAPES = [(16, 16, 16, False), (16, 32, 32, False), (32, 16, 64, True), (32, 32, 16, False), (32, 64, 16, False),
          (64, 32, 32, True), (64, 64, 16, False), (64, 64, 64, False), (16, 32, 64, False), (32, 32, 64, False),
          (64, 32, 64, True), (16, 16, 128, False), (32, 16, 128, True)]


@pytest.mark.parametrize("cin,cout,D,split", SHAPES)
def test_conv3_fwd(gpu, cin, cout, D, split):
    lib = gpu
    B = 1 if D >= 128 else 2
    x = _bf((B, D, D, D, cin), 0.7, 1)
    w = _bf((cout, 27, cin), np.sqrt(2.0 / (27 * cin)), 2)
    bias = _f((B, cout) if split else (cout,), 0.1, 3)
    c0 = cin // 2 if split else cin
    o0 = cout // 2 if split and cout >= 32 else cout
    x0, x1 = x[..., :c0].contiguous(), (x[..., c0:].contiguous() if split else None)
    y0 = torch.empty((B, D, D, D, o0), dtype=torch.bfloat16, device="cuda")
    y1 = torch.empty((B, D, D, D, cout - o0), dtype=torch.bfloat16, device="cuda") if o0 < cout else None
    ref = []
    for b in range(B):
        bb = _d(bias[b]) if split else _d(bias)
        ref.append(unet.conv3(_d(x[b:b + 1]), _d(w), bb)[0])
    ref = torch.stack(ref)
    for algo in ((1, 2, 3) if D >= 64 else (0, 3)):
        print(f"conv fwd cin={cin} cout={cout} D={D} split={split} algo={algo}", flush=True)
        y0.fill_(float("nan"))
        _conv_call(lib, B, D, cin, cout, x0, w, y0, x1=x1, c0=c0, bias=bias, y1=y1, o0=o0, per_sample=split,
                   algo=algo)
        torch.cuda.synchronize()
        _close_bf16(y0, ref[..., :o0], f"y0 algo {algo}")
        if y1 is not None:
            _close_bf16(y1, ref[..., o0:], f"y1 algo {algo}")



@pytest.mark.parametrize("cin,cout,D", [(16, 16, 32), (32, 16, 64), (32, 32, 64)])
def test_conv3_fwd_chunk_across_inputs(gpu, cin, cout, D):
    """Input split at c0 = 8: the 16-channel chunks of the chunked tile take
    channels from both inputs (the general load path; at c0 a multiple of 16
    every chunk lies in one input), also with the ReLU-mask input of the
    input-gradient form on a single input."""
    lib = gpu
    B = 2
    c0 = 8
    x = _bf((B, D, D, D, cin), 0.7, 11)
    w = _bf((cout, 27, cin), np.sqrt(2.0 / (27 * cin)), 12)
    bias = _f((cout,), 0.1, 13)
    x0, x1 = x[..., :c0].contiguous(), x[..., c0:].contiguous()
    y = torch.empty((B, D, D, D, cout), dtype=torch.bfloat16, device="cuda")
    ref = torch.stack([unet.conv3(_d(x[b:b + 1]), _d(w), _d(bias))[0] for b in range(B)])
    for algo in ((1, 2, 3) if D >= 64 else (0, 3)):
        print(f"conv fwd split 8 cin={cin} cout={cout} D={D} algo={algo}", flush=True)
        y.fill_(float("nan"))
        _conv_call(lib, B, D, cin, cout, x0, w, y, x1=x1, c0=c0, bias=bias, algo=algo)
        torch.cuda.synchronize()
        _close_bf16(y, ref, f"split 8 algo {algo}")
    # mask on a single input: both load paths give the same bits
    m = _bf((B, D, D, D, cin), 1.0, 14)
    outs = []
    for algo in ((1, 3) if D >= 64 else (0, 3)):
        yy = torch.empty_like(y)
        _conv_call(lib, B, D, cin, cout, x, w, yy, mask=m, relu=False, algo=algo)
        outs.append(yy)
    torch.cuda.synchronize()
    keep = (_d(m) > 0).to(torch.float64)
    refm = torch.stack([unet.conv3(_d(x[b:b + 1]) * keep[b:b + 1], _d(w), None, relu=False)[0] for b in range(B)])
    for yy in outs:
        _close_bf16(yy, refm, "masked input")

# weight gradient: fp32 VALU form at D = 16, matrix-core form at D >= 32
# (split: the input arrives as two tensors of cin/2 channels)
@pytest.mark.parametrize("cin,cout,D,split", [(16, 16, 32, False), (32, 16, 16, False), (64, 32, 16, False),
                                              (32, 64, 16, False), (16, 16, 64, False), (16, 32, 64, False),
                                              (32, 16, 64, True), (64, 32, 32, True), (32, 64, 32, False),
                                              (64, 64, 32, False)])
def test_conv3_input_and_weight_gradients(gpu, cin, cout, D, split):
    """Input gradient = the MFMA conv on the masked output gradient with the
    packed transposed weights; weight/bias gradients from the wgrad kernel;
    both against autograd of the oracle conv."""
    lib = gpu
    B = 2
    x = _bf((B, D, D, D, cin), 0.7, 11)
    w32 = _f((cout, 27, cin), np.sqrt(2.0 / (27 * cin)), 12)
    g = _bf((B, D, D, D, cout), 0.5, 13)
    m = _bf((B, D, D, D, cout), 1.0, 14)  # mixed signs: the ReLU mask
    wb = torch.empty((cout, 27, cin), dtype=torch.bfloat16, device="cuda")
    wt = torch.empty((cin, 27, cout), dtype=torch.bfloat16, device="cuda")
    lib.call("cm21_unet_pack", _fp(w32), _vp(wb), _vp(wt), cin, cout, lib.stream_handle())
    dx = torch.empty((B, D, D, D, cin), dtype=torch.bfloat16, device="cuda")
    _conv_call(lib, B, D, cout, cin, g, wt, dx, mask=m, relu=False)
    dw = torch.empty((cout, 27, cin), dtype=torch.float32, device="cuda")
    db = torch.empty(cout, dtype=torch.float32, device="cuda")
    work = torch.empty(lib.query("cm21_unet_wgrad_workspace_bytes", B, D, cin, cout), dtype=torch.uint8,
                       device="cuda")
    c0 = cin // 2 if split else cin
    xa, xb = (x[..., :c0].contiguous(), x[..., c0:].contiguous()) if split else (x, None)
    print(f"grad cin={cin} cout={cout} D={D} split={split}", flush=True)
    wg = lib.UnetWgrad(B=B, D=D, cin=cin, cout=cout, c0=c0, x0=_vp(xa), x1=_vp(xb), g=_vp(g), mask=_vp(m),
                       dw=_fp(dw), db=_fp(db), work=_vp(work), work_bytes=work.numel())
    lib.call("cm21_unet_conv3_wgrad", ctypes.byref(wg), lib.stream_handle())
    torch.cuda.synchronize()
    assert torch.equal(wb.cpu(), w32.cpu().to(torch.bfloat16))
    xr = _d(x).requires_grad_(True)
    wr = _d(wb).requires_grad_(True)
    br = torch.zeros(cout, dtype=torch.float64, requires_grad=True)
    gm = _d(g) * (_d(m) > 0)
    (unet.conv3(xr, wr, br, relu=False) * gm).sum().backward()
    _close_bf16(dx, xr.grad, "dx")
    # the weight gradient uses the bf16 input exactly as stored
    _close_f32(dw, wr.grad, "dw")
    _close_f32(db, br.grad, "db")


def test_pool_and_up(gpu):
    lib = gpu
    B, D = 2, 32
    x = _bf((B, D, D, D, 32), 1.0, 21)
    y = torch.empty((B, D // 2, D // 2, D // 2, 32), dtype=torch.bfloat16, device="cuda")
    lib.call("cm21_unet_pool_fwd", _vp(x), _vp(y), B, D, 32, lib.stream_handle())
    dy = _bf((B, D // 2, D // 2, D // 2, 32), 1.0, 22)
    dx = _bf((B, D, D, D, 32), 1.0, 23)
    dx0 = _d(dx)
    lib.call("cm21_unet_pool_bwd", _vp(dy), _vp(dx), B, D, 32, 1, lib.stream_handle())
    torch.cuda.synchronize()
    _close_bf16(y, unet.pool(_d(x)), "pool")
    up_dy = _d(dy).repeat_interleave(2, 1).repeat_interleave(2, 2).repeat_interleave(2, 3) / 8
    _close_bf16(dx, dx0 + up_dy, "pool bwd (accumulate)")
    for cin, cout, Din in ((64, 32, 16), (32, 16, 32)):
        xi = _bf((B, Din, Din, Din, cin), 0.8, 24)
        w = _f((cout, 8, cin), np.sqrt(1.0 / cin), 25)
        bia = _f((cout,), 0.1, 26)
        wn = torch.empty((8 * cout, cin), dtype=torch.bfloat16, device="cuda")
        wk = torch.empty((cin, 8 * cout), dtype=torch.bfloat16, device="cuda")
        lib.call("cm21_unet_pack_up", _fp(w), _vp(wn), _vp(wk), cin, cout, lib.stream_handle())
        yo = torch.empty((B, 2 * Din, 2 * Din, 2 * Din, cout), dtype=torch.bfloat16, device="cuda")
        lib.call("cm21_unet_up_fwd", _vp(xi), _vp(wn), _fp(bia), _vp(yo), B, Din, cin, cout, lib.stream_handle())
        g = _bf((B, 2 * Din, 2 * Din, 2 * Din, cout), 0.5, 27)
        gx = torch.empty_like(xi)
        dw = torch.empty_like(w)
        db = torch.empty_like(bia)
        work = torch.empty(lib.query("cm21_unet_up_workspace_bytes", B, Din, cin, cout), dtype=torch.uint8,
                           device="cuda")
        lib.call("cm21_unet_up_bwd", _vp(xi), _vp(g), _vp(wk), _vp(gx), _fp(dw), _fp(db), B, Din, cin, cout,
                 _vp(work), work.numel(), lib.stream_handle())
        torch.cuda.synchronize()
        # the matrix cores use the bf16 copy of the weights
        assert torch.equal(wn.cpu().float(), w.cpu().permute(1, 0, 2).reshape(8 * cout, cin).to(torch.bfloat16).float())
        xr = _d(xi).requires_grad_(True)
        wr = _d(w.to(torch.bfloat16)).requires_grad_(True)
        br = _d(bia).requires_grad_(True)
        out = unet.up(xr, wr, br)
        _close_bf16(yo, out.detach(), f"up fwd {cin}->{cout}")
        (out * _d(g)).sum().backward()
        _close_bf16(gx, xr.grad, f"up dx {cin}->{cout}")
        _close_f32(dw, wr.grad, f"up dw {cin}->{cout}")
        _close_f32(db, br.grad, f"up db {cin}->{cout}")


@pytest.mark.parametrize("D", [32, 64])
def test_input_layer(gpu, D):
    lib = gpu
    B = 2
    delta = _f((B, D, D, D), 0.6, 31)
    cond = torch.tensor([[0.3, -0.7], [-1.0, 0.9]], dtype=torch.float32, device="cuda")
    w = _f((16, 27, 3), np.sqrt(2.0 / 81), 32)
    bia = _f((16,), 0.1, 33)
    beff = torch.empty((B, 16), dtype=torch.float32, device="cuda")
    y = torch.empty((B, D, D, D, 16), dtype=torch.bfloat16, device="cuda")
    scale = 1 / 0.6
    lib.call("cm21_unet_in_fwd", _fp(delta), _fp(cond), _fp(w), _fp(bia), _fp(beff), _vp(y), B, D,
             ctypes.c_float(scale), lib.stream_handle())
    g = _bf((B, D, D, D, 16), 0.5, 34)
    dw = torch.empty_like(w)
    db = torch.empty_like(bia)
    work = torch.empty(lib.query("cm21_unet_in_workspace_bytes", B, D), dtype=torch.uint8, device="cuda")
    lib.call("cm21_unet_in_wgrad", _fp(delta), _fp(cond), _vp(g), _vp(y), _fp(dw), _fp(db), B, D,
             ctypes.c_float(scale), _vp(work), work.numel(), lib.stream_handle())
    torch.cuda.synchronize()
    dr = _d(delta)
    cr = _d(cond)
    x = torch.stack([dr * scale, cr[:, 0].reshape(B, 1, 1, 1).expand(B, D, D, D),
                     cr[:, 1].reshape(B, 1, 1, 1).expand(B, D, D, D)], -1)
    wr = _d(w).requires_grad_(True)
    br = _d(bia).requires_grad_(True)
    out = unet.conv3(x, wr, br)
    _close_bf16(y, out.detach(), "input layer")
    (out * (_d(g) * (_d(y) > 0))).sum().backward()
    _close_f32(dw, wr.grad, "input dw")
    _close_f32(db, br.grad, "input db")


def test_head_loss_and_gradients(gpu):
    lib = gpu
    B, D = 2, 32
    x = _bf((B, D, D, D, 16), 0.8, 41)
    w = _f((17,), 0.4, 42)
    target = torch.rand((B, D, D, D), generator=torch.Generator().manual_seed(43)).float().cuda()
    prob = torch.empty((B, D, D, D), dtype=torch.float32, device="cuda")
    dx = torch.empty_like(x)
    dwb = torch.empty(17, dtype=torch.float32, device="cuda")
    loss = torch.zeros(1, dtype=torch.float64, device="cuda")
    work = torch.empty(lib.query("cm21_unet_head_workspace_bytes"), dtype=torch.uint8, device="cuda")
    lib.call("cm21_unet_head", _vp(x), _fp(w[:16]), _fp(w[16:]), _fp(target), _fp(prob), _vp(dx), _fp(dwb),
             ctypes.cast(ctypes.c_void_p(loss.data_ptr()), ctypes.POINTER(ctypes.c_double)), B, D, _vp(work),
             work.numel(), lib.stream_handle())
    torch.cuda.synchronize()
    xr = _d(x).requires_grad_(True)
    wr = _d(w).requires_grad_(True)
    lg = (xr * wr[:16]).sum(-1) + wr[16]
    ref = unet.bce(lg, _d(target))
    ref.backward()
    assert float(loss) == pytest.approx(float(ref), rel=1e-6)
    _close_f32(prob, torch.sigmoid(lg).detach(), "prob", rel=1e-6)
    _close_bf16(dx, xr.grad, "head dx")
    _close_f32(dwb, wr.grad, "head dw/db")


def test_adamw(gpu):
    lib = gpu
    n = 1000
    rng = np.random.default_rng(5)
    p0, g, m0, v0 = (rng.standard_normal(n).astype(np.float32) for _ in range(4))
    v0 = np.abs(v0)
    p, m, v = (torch.from_numpy(a.copy()).cuda() for a in (p0, m0, v0))
    gt = torch.from_numpy(g).cuda()
    lr, b1, b2, eps, wd, step = 1e-3, 0.9, 0.999, 1e-8, 0.01, 7
    lib.call("cm21_adamw", _fp(p), _fp(gt), _fp(m), _fp(v), n, ctypes.c_float(lr), ctypes.c_float(b1),
             ctypes.c_float(b2), ctypes.c_float(eps), ctypes.c_float(wd), step, lib.stream_handle())
    torch.cuda.synchronize()
    mm = b1 * m0.astype(np.float64) + (1 - b1) * g
    vv = b2 * v0.astype(np.float64) + (1 - b2) * g.astype(np.float64) ** 2
    pp = p0 - lr * ((mm / (1 - b1 ** step)) / (np.sqrt(vv / (1 - b2 ** step)) + eps) + wd * p0)
    np.testing.assert_allclose(m.cpu().numpy(), mm, rtol=1e-5, atol=1e-7)
    np.testing.assert_allclose(v.cpu().numpy(), vv, rtol=1e-5, atol=1e-7)
    np.testing.assert_allclose(p.cpu().numpy(), pp, rtol=1e-5, atol=1e-6)


def _net_inputs(D, B, seed):
    rng = np.random.default_rng(seed)
    delta = (rng.standard_normal((B, D, D, D)) * 0.6).astype(np.float32)
    cond = rng.uniform(-1, 1, (B, 2)).astype(np.float32)
    target = rng.uniform(0, 1, (B, D, D, D)).astype(np.float32)
    target[:, : D // 2] = (target[:, : D // 2] > 0.5)
    return delta, cond, target


def test_network_forward_and_gradients_match_oracle(gpu):
    sur = pkg("surrogate")
    D, B = 64, 2
    P = sur.init_params(1)
    rng = np.random.default_rng(9)
    for k in P:
        if k.endswith(".b"):
            P[k] = rng.standard_normal(P[k].shape) * 0.05
    model = sur.UNet3D(D=D, batch=B, params=P)
    # the GPU keeps fp32 masters: compare against the oracle on the same fp32 values
    P32 = {k: v.astype(np.float32).astype(np.float64) for k, v in P.items()}
    delta, cond, target = _net_inputs(D, B, 10)
    prob = model.forward(torch.from_numpy(delta).cuda(), cond).cpu().numpy()
    loss = float(model.backward(torch.from_numpy(target).cuda()))
    torch.cuda.synchronize()
    ref_prob = unet.predict(P32, delta, cond, sur.DELTA_SCALE, bf16=True)
    err = np.abs(prob - ref_prob)
    print(f"prob err mean {err.mean():.2e} max {err.max():.2e}", flush=True)
    assert err.mean() < 2e-3 and err.max() < 5e-2
    ref_loss, ref_g = unet.loss_and_grads(P32, delta, cond, target, sur.DELTA_SCALE, bf16=True)
    assert loss == pytest.approx(ref_loss, rel=2e-3)
    for name, gref in ref_g.items():
        gg = model.g[name].detach().cpu().double().numpy()
        rel = np.linalg.norm(gg - gref) / max(np.linalg.norm(gref), 1e-30)
        cos = float(np.dot(gg.ravel(), gref.ravel()) / max(np.linalg.norm(gg) * np.linalg.norm(gref), 1e-30))
        print(f"{name}: rel L2 {rel:.2e} cos {cos:.5f}", flush=True)
        assert rel < 0.1 and cos > 0.99, name


def test_training_reduces_loss_on_plain_targets(gpu):
    """A few AdamW steps on boxes labelled by the plain method lower the loss."""
    sur = pkg("surrogate")
    specs = sur.surrogate_specs(range(1000, 1008), D=64, box_len=100.0)
    boxes = sur.BoxSet(specs)
    xhi = boxes.x_HI()
    print("x_HI of the boxes", np.round(xhi, 3), flush=True)
    assert np.any((xhi > 0.05) & (xhi < 0.95))
    model = sur.UNet3D(D=64, batch=2, seed=0)
    losses = sur.train(model, boxes, steps=40, lr=3e-3, log_every=1)
    first = np.mean([l for _, l in losses[:5]])
    last = np.mean([l for _, l in losses[-5:]])
    print(f"loss {first:.4f} -> {last:.4f}", flush=True)
    assert last < 0.8 * first


def _heldout(n, D=128, first=3000):
    sur = pkg("surrogate")
    return sur.BoxSet(sur.surrogate_specs(range(first, first + n), seed=8, D=D, box_len=200.0))


def test_trained_network_gains_over_untrained(gpu):
    """The committed weights (trained on seeds 1000..2999 and 3200..7199) against an
    untrained network of the same shape on held-out seeds: the trained one
    reproduces the plain method's neutral fraction and ionized cells, the
    untrained one does not."""
    sur = pkg("surrogate")
    params, meta = sur.UNet3D.load_file(sur.WEIGHTS)
    boxes = _heldout(8)
    trained = sur.summarise(sur.evaluate(sur.UNet3D(D=128, batch=4, params=params), boxes))
    untrained = sur.summarise(sur.evaluate(sur.UNet3D(D=128, batch=4, seed=123), boxes))
    print("trained", trained, "\nuntrained", untrained, flush=True)
    assert trained["dx_HI_mean"] < 0.02 and trained["agreement_mean"] > 0.95
    assert trained["dx_HI_mean"] < 0.25 * untrained["dx_HI_mean"]
    assert trained["agreement_mean"] > untrained["agreement_mean"] + 0.1


def test_pipeline_validates_and_falls_back(gpu):
    """Served with the trained network at the production tolerances (|dx_HI| <
    0.01, P21 within 10 %): the validated box decides the answer.  The
    network's prediction and its metrics against the plain method are computed
    independently first; a pipeline whose sampled box meets the tolerances
    answers with the network's fields, one whose sampled box misses them
    answers the whole request with the plain fields.  An untrained network
    fails validation, and after max_failures it is bypassed."""
    sur = pkg("surrogate")
    params, _ = sur.UNet3D.load_file(sur.WEIGHTS)
    boxes = _heldout(8)
    model = sur.UNet3D(D=128, batch=4, params=params)
    n_surrogate = n_plain = 0
    for first in (0, 4):
        delta = boxes.delta[first:first + 4].contiguous()
        zeta = [s["zeta"] for s in boxes.specs[first:first + 4]]
        smin = [s["sigma_min2"] for s in boxes.specs[first:first + 4]]
        cond = [sur.normalise_cond(z, s) for z, s in zip(zeta, smin)]
        prob = model.forward(delta, cond, keep=False).clone()
        metrics = [sur.box_metrics(prob[j], boxes.target[first + j], delta[j], 200.0) for j in range(4)]
        for j in range(4):
            m = metrics[j]
            ok = m["dx_HI"] <= 0.01 and m["p21_err"] <= 0.10
            print(f"box {first + j}: dx_HI {m['dx_HI']:.4f} P21 {m['p21_err']:.3f} -> "
                  f"{'within' if ok else 'outside'} the tolerances", flush=True)
            pipe = sur.SurrogatePipeline(model, validate_every=4)  # production tolerances
            assert (pipe.tol_xhi, pipe.tol_p21) == (0.01, 0.10)
            pipe.seen = (4 - j) % 4  # the request's box j is the sampled one
            fields, src = pipe(delta, zeta, smin)
            assert pipe.validated == 1
            if ok:
                n_surrogate += 1
                assert src == "surrogate" and pipe.fallbacks == 0
                assert torch.equal(fields, prob)
            else:
                n_plain += 1
                assert src == "plain" and pipe.fallbacks == 1 and pipe.failures == 1
                for i in range(4):
                    assert torch.equal(fields[i], boxes.target[first + i]), (first, j, i)
    print(f"answered by the network {n_surrogate}, by the plain method {n_plain}", flush=True)
    assert n_surrogate > 0  # the trained network passes on some held-out boxes
    delta = boxes.delta[:4].contiguous()
    zeta = [s["zeta"] for s in boxes.specs[:4]]
    smin = [s["sigma_min2"] for s in boxes.specs[:4]]
    bad = sur.SurrogatePipeline(sur.UNet3D(D=128, batch=4, seed=5), validate_every=1, max_failures=2)
    for rep in range(3):
        fields, src = bad(delta, zeta, smin)
        assert src == "plain"
        for j in range(4):
            assert torch.equal(fields[j], boxes.target[j]), (rep, j)
    assert bad.failures == 2 and bad.disabled and bad.fallbacks == 3


@pytest.mark.parametrize("D", [64, 128])
def test_fused_head_matches_unfused(gpu, D):
    """Inference fuses the head into the last convolution: with the same conv
    form the probabilities equal those of the separate head kernel bit for bit
    (both sum the logit in the same order); the model's fused and stored
    passes agree to bf16 rounding; backward() refuses to run on the
    activations the fused pass did not store."""
    sur = pkg("surrogate")
    B = 1
    P = sur.init_params(3)
    rng = np.random.default_rng(4)
    for k in P:
        if k.endswith(".b"):
            P[k] = rng.standard_normal(P[k].shape) * 0.05
    model = sur.UNet3D(D=D, batch=B, params=P)
    delta, cond, _ = _net_inputs(D, B, 12)
    d = torch.from_numpy(delta).cuda()
    p_sep = model.forward(d, cond).clone()
    p_fused = model.forward(d, cond, keep=False).clone()
    torch.cuda.synchronize()
    diff = (p_sep - p_fused).abs()
    print(f"D={D}: max |diff| {float(diff.max()):.3g} mean {float(diff.mean()):.3g}", flush=True)
    # the stored-activation pass and the fused pass may run the last conv in
    # different forms (K summed in another order): a bf16 rounding of an
    # activation may flip, nothing more
    assert float(diff.max()) < 1e-2 and float(diff.mean()) < 1e-5
    with pytest.raises(RuntimeError):
        model.backward(torch.zeros_like(d))
    # every conv form that carries the head against the same form's stored
    # output through the separate head kernel (the forms accumulate K in
    # different orders, so each is compared with itself)
    lib = gpu
    a, p = model.a, model.p
    model.forward(d, cond)
    for algo in ((1, 2, 3) if D >= 64 else (0, 3)):
        y = torch.empty_like(a["d1"])
        ref = torch.full_like(model.prob, float("nan"))
        prob = torch.full_like(model.prob, float("nan"))
        common = dict(B=B, D=D, cin=16, cout=16, c0=16, o0=16, x0=_vp(a["d1a"]), w=_vp(model.wb["d1b"]),
                      bias=_fp(p["d1b.b"]), relu=1, algo=algo)
        c = lib.UnetConv(y0=_vp(y), **common)
        lib.call("cm21_unet_conv3_fwd", ctypes.byref(c), lib.stream_handle())
        lib.call("cm21_unet_head", _vp(y), _fp(p["head.w"]), _fp(p["head.b"]), None, _fp(ref), None, None, None,
                 B, D, None, 0, lib.stream_handle())
        c = lib.UnetConv(y0=None, head=_fp(p["head.w"]), prob=_fp(prob), **common)
        lib.call("cm21_unet_conv3_fwd", ctypes.byref(c), lib.stream_handle())
        torch.cuda.synchronize()
        print(f"algo {algo}: max |diff| {float((prob - ref).abs().max()):.3g}", flush=True)
        assert torch.equal(prob, ref), f"algo {algo}"


@pytest.mark.parametrize("D", [64, 128])
def test_conv_pool_output_matches_pool_kernel(gpu, D):
    """A conv given a pool buffer also writes the 2x2x2 mean of its output:
    bit for bit what cm21_unet_pool_fwd makes of the stored output, for every
    conv form (the sliding form of 16 -> 16 channels computes it in its
    epilogue, the others pool afterwards), with the stored output unchanged;
    with and without the ReLU."""
    lib = gpu
    B = 1 if D >= 128 else 2
    st = lib.stream_handle()
    for cin, cout in ((16, 16), (32, 32)):
        x = _bf((B, D, D, D, cin), 1.0, 41)
        w = _bf((cout, 27, cin), np.sqrt(2.0 / (27 * cin)), 42)
        bias = _f((cout,), 0.1, 43)
        for algo in (1, 2, 3):
            for relu in ((True, False) if (cin, algo) == (16, 2) else (True,)):
                print(f"pool fused: {cin}->{cout} D={D} algo={algo} relu={relu}", flush=True)
                y_ref = torch.empty((B, D, D, D, cout), dtype=torch.bfloat16, device="cuda")
                _conv_call(lib, B, D, cin, cout, x, w, y_ref, bias=bias, relu=relu, algo=algo)
                p_ref = torch.empty((B, D // 2, D // 2, D // 2, cout), dtype=torch.bfloat16, device="cuda")
                lib.call("cm21_unet_pool_fwd", _vp(y_ref), _vp(p_ref), B, D, cout, st)
                y = torch.full_like(y_ref, float("nan"))
                pooled = torch.full_like(p_ref, float("nan"))
                c = lib.UnetConv(B=B, D=D, cin=cin, cout=cout, c0=cin, o0=cout, x0=_vp(x), w=_vp(w), bias=_fp(bias),
                                 relu=1 if relu else 0, y0=_vp(y), algo=algo, pool=_vp(pooled))
                lib.call("cm21_unet_conv3_fwd", ctypes.byref(c), st)
                torch.cuda.synchronize()
                assert torch.equal(y.view(torch.int16), y_ref.view(torch.int16)), f"y algo {algo}"
                assert torch.equal(pooled.view(torch.int16), p_ref.view(torch.int16)), f"pool algo {algo}"
    # the pool needs every output channel in y0 and no head
    c = lib.UnetConv(B=B, D=D, cin=16, cout=16, c0=16, o0=16, x0=_vp(x), w=_vp(w), y0=_vp(y), pool=_vp(pooled),
                     head=_fp(bias), prob=_fp(bias))
    with pytest.raises(RuntimeError):
        lib.call("cm21_unet_conv3_fwd", ctypes.byref(c), st)
