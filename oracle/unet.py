"""CPU restatement of the 3D U-Net surrogate (TEST INFRASTRUCTURE ONLY).

The surrogate has no counterpart in the reference; this module states its
layers with torch CPU float64 operators (conv3d on circularly padded input,
conv_transpose3d, avg_pool3d, sigmoid + binary cross entropy) so that the GPU
layer kernels and the whole forward/backward pass can be checked against it.
Gradients come from torch autograd in float64.

Layouts follow the GPU side: activations (B, D, D, D, C) channels-last;
3^3 conv weights [cout][27][cin] with tap t = 9 (dz+1) + 3 (dy+1) + (dx+1)
and y[v] += w[t] x[v + (dz, dy, dx)] (periodic); transposed-conv weights
[cout][8][cin] with sub-voxel s = 4a + 2b + c placing x[v] at 2v + (a, b, c).
With bf16=True every tensor the GPU stores in bf16 (3^3 conv weights, layer
outputs) is rounded to bf16 (round to nearest even) in the forward pass.
"""
import numpy as np
import torch
import torch.nn.functional as F

# name, kind, cin, cout (input-layer cin counts the density and the two
# conditioning channels)
LAYERS = [
    ("in", "in", 3, 16),
    ("e1", "conv", 16, 16),
    ("e2a", "conv", 16, 32),
    ("e2b", "conv", 32, 32),
    ("b1", "conv", 32, 64),
    ("b2", "conv", 64, 64),
    ("up2", "up", 64, 32),
    ("d2a", "conv", 64, 32),
    ("d2b", "conv", 32, 32),
    ("up1", "up", 32, 16),
    ("d1a", "conv", 32, 16),
    ("d1b", "conv", 16, 16),
    ("head", "head", 16, 1),
]


def bf16_round(x):
    """Round float tensor values to bf16 (nearest even), returned in float64."""
    return x.to(torch.float32).to(torch.bfloat16).to(torch.float64)


class _RoundBF16(torch.autograd.Function):
    @staticmethod
    def forward(ctx, x):
        return bf16_round(x)

    @staticmethod
    def backward(ctx, g):
        return g


def _r(x, on):
    return _RoundBF16.apply(x) if on else x


def to_ncdhw(x):
    """(B, D, D, D, C) -> (B, C, D, D, D)."""
    return x.permute(0, 4, 1, 2, 3)


def to_ndhwc(x):
    return x.permute(0, 2, 3, 4, 1)


def conv3(x, w, b=None, relu=True):
    """x (B,D,D,D,cin), w [cout][27][cin] -> (B,D,D,D,cout), periodic."""
    cout, _, cin = w.shape
    wt = w.reshape(cout, 3, 3, 3, cin).permute(0, 4, 1, 2, 3)
    xp = F.pad(to_ncdhw(x), (1, 1, 1, 1, 1, 1), mode="circular")
    y = F.conv3d(xp, wt, b)
    if relu:
        y = F.relu(y)
    return to_ndhwc(y)


def up(x, w, b):
    """x (B,D,D,D,cin), w [cout][8][cin] -> (B,2D,2D,2D,cout)."""
    cout, _, cin = w.shape
    wt = w.reshape(cout, 2, 2, 2, cin).permute(4, 0, 1, 2, 3)
    return to_ndhwc(F.conv_transpose3d(to_ncdhw(x), wt, b, stride=2))


def pool(x):
    return to_ndhwc(F.avg_pool3d(to_ncdhw(x), 2))


def init_params(seed=0):
    """He-normal weights, zero biases (float64 numpy arrays)."""
    rng = np.random.default_rng(seed)
    p = {}
    for name, kind, cin, cout in LAYERS:
        if kind in ("in", "conv"):
            p[name + ".w"] = rng.standard_normal((cout, 27, cin)) * np.sqrt(2.0 / (27 * cin))
            p[name + ".b"] = np.zeros(cout)
        elif kind == "up":
            p[name + ".w"] = rng.standard_normal((cout, 8, cin)) * np.sqrt(1.0 / cin)
            p[name + ".b"] = np.zeros(cout)
        else:
            p[name + ".w"] = rng.standard_normal(cin) * np.sqrt(1.0 / cin)
            p[name + ".b"] = np.zeros(1)
    return p


def forward(params, delta, cond, delta_scale, bf16=False):
    """Logits (B, D, D, D) of the network.  params: dict of float64 tensors;
    delta (B,D,D,D); cond (B,2)."""
    P = params
    B, D = delta.shape[0], delta.shape[1]
    x = torch.stack([delta * delta_scale,
                     cond[:, 0].reshape(B, 1, 1, 1).expand(B, D, D, D),
                     cond[:, 1].reshape(B, 1, 1, 1).expand(B, D, D, D)], dim=-1)

    def cw(name):
        return _r(P[name + ".w"], bf16)

    e1a = _r(conv3(x, P["in.w"], P["in.b"]), bf16)
    e1 = _r(conv3(e1a, cw("e1"), P["e1.b"]), bf16)
    p1 = _r(pool(e1), bf16)
    e2a = _r(conv3(p1, cw("e2a"), P["e2a.b"]), bf16)
    e2 = _r(conv3(e2a, cw("e2b"), P["e2b.b"]), bf16)
    p2 = _r(pool(e2), bf16)
    b1 = _r(conv3(p2, cw("b1"), P["b1.b"]), bf16)
    b2 = _r(conv3(b1, cw("b2"), P["b2.b"]), bf16)
    u2 = _r(up(b2, P["up2.w"], P["up2.b"]), bf16)
    d2a = _r(conv3(torch.cat([e2, u2], -1), cw("d2a"), P["d2a.b"]), bf16)
    d2 = _r(conv3(d2a, cw("d2b"), P["d2b.b"]), bf16)
    u1 = _r(up(d2, P["up1.w"], P["up1.b"]), bf16)
    d1a = _r(conv3(torch.cat([e1, u1], -1), cw("d1a"), P["d1a.b"]), bf16)
    d1 = _r(conv3(d1a, cw("d1b"), P["d1b.b"]), bf16)
    return (d1 * P["head.w"]).sum(-1) + P["head.b"][0]


def bce(logits, target):
    """Mean binary cross entropy of sigmoid(logits) against target in [0,1]."""
    return F.binary_cross_entropy_with_logits(logits, target, reduction="mean")


def loss_and_grads(params_np, delta, cond, target, delta_scale, bf16=False):
    """(loss, {name: grad}) in float64 for numpy inputs."""
    P = {k: torch.tensor(v, dtype=torch.float64, requires_grad=True) for k, v in params_np.items()}
    d = torch.tensor(np.asarray(delta), dtype=torch.float64)
    c = torch.tensor(np.asarray(cond), dtype=torch.float64)
    t = torch.tensor(np.asarray(target), dtype=torch.float64)
    loss = bce(forward(P, d, c, delta_scale, bf16), t)
    loss.backward()
    return float(loss.detach()), {k: v.grad.numpy() for k, v in P.items()}


def predict(params_np, delta, cond, delta_scale, bf16=False):
    P = {k: torch.tensor(v, dtype=torch.float64) for k, v in params_np.items()}
    with torch.no_grad():
        lg = forward(P, torch.tensor(np.asarray(delta), dtype=torch.float64),
                     torch.tensor(np.asarray(cond), dtype=torch.float64), delta_scale, bf16)
    return torch.sigmoid(lg).numpy()
