"""CPU pinning of the U-Net oracle (oracle/unet.py): each layer against a
brute-force loop over voxels and taps, the loss against its closed form, and
the autograd gradients of the whole network against central finite
differences."""
import numpy as np
import pytest
import torch

from conftest import pkg
from oracle import unet


def _loop_conv3(x, w, b, relu=True):
    B, D = x.shape[0], x.shape[1]
    cout = w.shape[0]
    y = np.zeros((B, D, D, D, cout))
    for bb in range(B):
        for z in range(D):
            for yy in range(D):
                for xx in range(D):
                    acc = b.copy()
                    for t in range(27):
                        dz, dy, dx = t // 9 - 1, (t // 3) % 3 - 1, t % 3 - 1
                        acc += w[:, t, :] @ x[bb, (z + dz) % D, (yy + dy) % D, (xx + dx) % D]
                    y[bb, z, yy, xx] = np.maximum(acc, 0) if relu else acc
    return y


def test_conv3_matches_loop():
    rng = np.random.default_rng(0)
    x = rng.standard_normal((2, 4, 4, 4, 3))
    w = rng.standard_normal((5, 27, 3))
    b = rng.standard_normal(5)
    got = unet.conv3(torch.tensor(x), torch.tensor(w), torch.tensor(b)).numpy()
    np.testing.assert_allclose(got, _loop_conv3(x, w, b), rtol=1e-12, atol=1e-12)
    got = unet.conv3(torch.tensor(x), torch.tensor(w), torch.tensor(b), relu=False).numpy()
    np.testing.assert_allclose(got, _loop_conv3(x, w, b, relu=False), rtol=1e-12, atol=1e-12)


def test_up_and_pool_match_loops():
    rng = np.random.default_rng(1)
    x = rng.standard_normal((1, 3, 3, 3, 4))
    w = rng.standard_normal((2, 8, 4))
    b = rng.standard_normal(2)
    got = unet.up(torch.tensor(x), torch.tensor(w), torch.tensor(b)).numpy()
    want = np.zeros((1, 6, 6, 6, 2))
    for z in range(3):
        for y in range(3):
            for xx in range(3):
                for s in range(8):
                    a, bb, c = s >> 2, (s >> 1) & 1, s & 1
                    want[0, 2 * z + a, 2 * y + bb, 2 * xx + c] = b + w[:, s, :] @ x[0, z, y, xx]
    np.testing.assert_allclose(got, want, rtol=1e-12, atol=1e-12)
    x = rng.standard_normal((2, 4, 4, 4, 3))
    got = unet.pool(torch.tensor(x)).numpy()
    want = x.reshape(2, 2, 2, 2, 2, 2, 2, 3).mean(axis=(2, 4, 6))
    np.testing.assert_allclose(got, want, rtol=1e-12, atol=1e-12)


def test_bce_closed_form():
    lg = torch.tensor([-30.0, -2.0, 0.0, 0.7, 25.0], dtype=torch.float64)
    t = torch.tensor([0.0, 0.3, 1.0, 0.5, 1.0], dtype=torch.float64)
    p = 1 / (1 + np.exp(-lg.numpy()))
    want = -np.mean(t.numpy() * np.log(p) + (1 - t.numpy()) * np.log1p(-p + 1e-300))
    assert float(unet.bce(lg, t)) == pytest.approx(want, rel=1e-9, abs=1e-12)


def test_bf16_round_is_nearest_even():
    v = torch.tensor([1.0 + 2 ** -8, 1.0 + 3 * 2 ** -8, 1.0 + 2 ** -9, -3.14159], dtype=torch.float64)
    got = unet.bf16_round(v).numpy()
    np.testing.assert_array_equal(got, [1.0, 1.0 + 2 ** -6, 1.0, -3.140625])


def test_network_gradients_match_finite_differences():
    """Whole network at D=4 (two poolings down to 1^3), float64: autograd
    gradients of a few entries of every parameter tensor against central
    differences."""
    rng = np.random.default_rng(2)
    P = unet.init_params(3)
    for k in P:
        if k.endswith(".b"):
            P[k] = rng.standard_normal(P[k].shape) * 0.1
    B, D = 2, 4
    delta = rng.standard_normal((B, D, D, D)) * 0.6
    cond = rng.uniform(-1, 1, (B, 2))
    target = rng.uniform(0, 1, (B, D, D, D))
    loss, grads = unet.loss_and_grads(P, delta, cond, target, 1 / 0.6)
    h = 1e-6
    for name in P:
        flat = P[name].reshape(-1)
        for idx in rng.choice(flat.size, size=min(3, flat.size), replace=False):
            old = flat[idx]
            flat[idx] = old + h
            lp = unet.loss_and_grads(P, delta, cond, target, 1 / 0.6)[0]
            flat[idx] = old - h
            lm = unet.loss_and_grads(P, delta, cond, target, 1 / 0.6)[0]
            flat[idx] = old
            fd = (lp - lm) / (2 * h)
            assert grads[name].reshape(-1)[idx] == pytest.approx(fd, rel=1e-5, abs=1e-9), (name, idx)


def test_surrogate_parameter_set_matches_oracle():
    sur = pkg("surrogate")
    assert [l[:4] for l in sur.LAYERS] == [l[:4] for l in unet.LAYERS]
    a, b = sur.init_params(5), unet.init_params(5)
    assert set(a) == set(b)
    for k in a:
        np.testing.assert_array_equal(a[k], b[k])
    n = sum(v.size for v in a.values())
    assert 300_000 < n < 400_000


def test_conditioning_normalisation():
    sur = pkg("surrogate")
    assert sur.normalise_cond(10.0, 2.0) == [-1.0, -1.0]
    assert sur.normalise_cond(60.0, 6.0) == [1.0, 1.0]
    assert sur.normalise_cond(35.0, 4.0) == [0.0, 0.0]
