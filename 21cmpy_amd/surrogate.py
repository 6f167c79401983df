"""3D U-Net surrogate of find_HII_bubbles, trained on the plain method.

The network maps a linear density box (HII_DIM^3, fp32) plus the ionizing
efficiency zeta and sigma_min^2 (two constant conditioning channels) to a
per-cell ionized probability.  Three levels with 16/32/64 channels, two 3^3
convolutions per level (ReLU), 2^3 mean pooling down, 2^3 stride-2
transposed convolutions up, skip connections by channel concatenation, a 1^3
head and a sigmoid.  Every layer runs in the HIP kernels of
csrc/cm21_unet.hip (3^3 convolutions and their input gradients on the bf16
matrix cores, fp32 accumulation); this module owns the buffers and the
order of the calls.  Training minimises the mean binary cross entropy
against the plain method's ionized field with AdamW.

``SurrogatePipeline`` serves requests with the network and checks a sample of
them against the plain method (the fallback and the reference answer): when a sampled
box misses the tolerance (|dx_HI| or the P21(k) error) the batch is
recomputed with the plain method.

The excursion-set reference (Programs/find_HII_bubbles.py:17-193) has no
learned component; the interfaces here are new.
"""
import ctypes
import importlib
import json
import math
import os

import numpy as np
import torch

from . import _lib

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

DELTA_SCALE = 1.0 / 0.6              # density in units of the per-cell sigma
ZETA_RANGE = (10.0, 60.0)
SMIN2_RANGE = (2.0, 6.0)
_PKG = os.path.dirname(os.path.abspath(__file__))
WEIGHTS = os.path.join(_PKG, "Parameter_files", "unet128_weights.npz")


def normalise_cond(zeta, sigma_min2):
    """Conditioning channels in [-1, 1] over the training ranges."""
    z0, z1 = ZETA_RANGE
    s0, s1 = SMIN2_RANGE
    return [(2.0 * zeta - z0 - z1) / (z1 - z0), (2.0 * sigma_min2 - s0 - s1) / (s1 - s0)]


def param_shapes():
    out = []
    for name, kind, cin, cout in LAYERS:
        if kind in ("in", "conv"):
            out += [(name + ".w", (cout, 27, cin)), (name + ".b", (cout,))]
        elif kind == "up":
            out += [(name + ".w", (cout, 8, cin)), (name + ".b", (cout,))]
        else:
            out += [(name + ".w", (cin,)), (name + ".b", (1,))]
    return out


def init_params(seed=0):
    """He-normal weights, zero biases (float64 numpy)."""
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


def _vp(t):
    return None if t is None else ctypes.c_void_p(t.data_ptr())


def _fp(t):
    return None if t is None else ctypes.cast(ctypes.c_void_p(t.data_ptr()), ctypes.POINTER(ctypes.c_float))


class UNet3D:
    """Parameters, optimiser state and activation buffers for batches of B
    boxes of side D (a multiple of 16)."""

    def __init__(self, D=128, batch=2, params=None, seed=0, stream=None):
        _lib.require_gpu()
        if D % 16:
            raise ValueError("D must be a multiple of 16")
        self.D, self.B = int(D), int(batch)
        self.stream = stream
        dev = torch.device("cuda")
        shapes = param_shapes()
        n = sum(int(np.prod(s)) for _, s in shapes)
        self.nparams = n
        self.flat = torch.zeros(n, dtype=torch.float32, device=dev)
        self.grad = torch.zeros_like(self.flat)
        self.m = torch.zeros_like(self.flat)
        self.v = torch.zeros_like(self.flat)
        self.p, self.g = {}, {}
        off = 0
        for name, s in shapes:
            k = int(np.prod(s))
            self.p[name] = self.flat[off:off + k].view(*s)
            self.g[name] = self.grad[off:off + k].view(*s)
            off += k
        self.load_params(params if params is not None else init_params(seed))
        self.wb, self.wt = {}, {}
        for name, kind, cin, cout in LAYERS:
            if kind == "conv":
                self.wb[name] = torch.empty((cout, 27, cin), dtype=torch.bfloat16, device=dev)
                self.wt[name] = torch.empty((cin, 27, cout), dtype=torch.bfloat16, device=dev)
            elif kind == "up":
                self.wb[name] = torch.empty((8 * cout, cin), dtype=torch.bfloat16, device=dev)
                self.wt[name] = torch.empty((cin, 8 * cout), dtype=torch.bfloat16, device=dev)
        self.step_count = 0
        self._alloc(dev)
        self.pack()

    # -------------------------------------------------------------- buffers
    def _alloc(self, dev):
        B, D = self.B, self.D
        h, q = D // 2, D // 4

        def act(side, c):
            return torch.empty((B, side, side, side, c), dtype=torch.bfloat16, device=dev)

        self.a = {"e1a": act(D, 16), "e1": act(D, 16), "p1": act(h, 16), "e2a": act(h, 32), "e2": act(h, 32),
                  "p2": act(q, 32), "b1": act(q, 64), "b2": act(q, 64), "u2": act(h, 32), "d2a": act(h, 32),
                  "d2": act(h, 32), "u1": act(D, 16), "d1a": act(D, 16), "d1": act(D, 16)}
        self.ga = {k: torch.empty_like(v) for k, v in self.a.items()}
        self.prob = torch.empty((B, D, D, D), dtype=torch.float32, device=dev)
        self.cond = torch.zeros((B, 2), dtype=torch.float32, device=dev)
        self.beff = torch.zeros((B, 16), dtype=torch.float32, device=dev)
        self.loss = torch.zeros(1, dtype=torch.float64, device=dev)
        need = max(_lib.query("cm21_unet_wgrad_workspace_bytes", B, s, ci, co)
                   for s, ci, co in [(D, 16, 16), (h, 16, 32), (h, 32, 32), (q, 32, 64), (q, 64, 64), (h, 64, 32),
                                     (D, 32, 16)])
        need = max(need, _lib.query("cm21_unet_up_workspace_bytes", B, q, 64, 32),
                   _lib.query("cm21_unet_up_workspace_bytes", B, h, 32, 16),
                   _lib.query("cm21_unet_in_workspace_bytes", B, D), _lib.query("cm21_unet_head_workspace_bytes"))
        self.work = torch.empty(need, dtype=torch.uint8, device=dev)
        self.delta = None

    def _st(self):
        return _lib.stream_handle(self.stream)

    # -------------------------------------------------------------- params
    def load_params(self, params):
        for name, _ in param_shapes():
            self.p[name].copy_(torch.as_tensor(np.asarray(params[name]), dtype=torch.float32))

    def params_numpy(self):
        return {k: v.detach().cpu().double().numpy() for k, v in self.p.items()}

    def save(self, path, **meta):
        np.savez(path, **{k: v.detach().cpu().numpy() for k, v in self.p.items()},
                 __meta__=np.array(json.dumps(meta)))

    def save_checkpoint(self, path, rng=None, **meta):
        """A training checkpoint: the fp32 master weights, the AdamW moments m
        and v, the step count (bias correction and schedule) and, given the
        training generator, its PCG64 state, so that a run resumed with
        load_checkpoint continues the same sequence of batches and updates."""
        state = {"step_count": int(self.step_count)}
        if rng is not None:
            state["rng"] = rng.bit_generator.state
        np.savez(path, **{k: v.detach().cpu().numpy() for k, v in self.p.items()},
                 __adam_m__=self.m.cpu().numpy(), __adam_v__=self.v.cpu().numpy(),
                 __state__=np.array(json.dumps(state)), __meta__=np.array(json.dumps(meta)))

    def load_checkpoint(self, path):
        """Restore weights, AdamW moments and step count from save_checkpoint;
        returns a NumPy Generator in the saved state (or None)."""
        with np.load(path, allow_pickle=False) as f:
            self.load_params({k: f[k] for k in f.files if not k.startswith("__")})
            self.m.copy_(torch.from_numpy(f["__adam_m__"]))
            self.v.copy_(torch.from_numpy(f["__adam_v__"]))
            state = json.loads(str(f["__state__"]))
        self.step_count = int(state["step_count"])
        self.pack()
        if "rng" not in state:
            return None
        rng = np.random.Generator(np.random.PCG64())
        rng.bit_generator.state = state["rng"]
        return rng

    @staticmethod
    def load_file(path):
        with np.load(path, allow_pickle=False) as f:
            params = {k: f[k].astype(np.float64) for k in f.files if not k.startswith("__")}
            meta = json.loads(str(f["__meta__"])) if "__meta__" in f.files else {}
        return params, meta

    def pack(self):
        """bf16 copies of the fp32 masters for the matrix-core kernels."""
        for name, kind, cin, cout in LAYERS:
            if kind == "conv":
                _lib.call("cm21_unet_pack", _fp(self.p[name + ".w"]), _vp(self.wb[name]), _vp(self.wt[name]), cin,
                          cout, self._st())
            elif kind == "up":
                _lib.call("cm21_unet_pack_up", _fp(self.p[name + ".w"]), _vp(self.wb[name]), _vp(self.wt[name]), cin,
                          cout, self._st())

    # -------------------------------------------------------------- layers
    def _conv(self, x0, w, y0, cin, cout, D, x1=None, mask=None, bias=None, relu=True, y1=None, c0=None, o0=None,
              bias_per_sample=False, head=None, prob=None, pool=None):
        c = _lib.UnetConv(B=self.B, D=D, cin=cin, cout=cout, c0=c0 or cin, o0=o0 or cout, x0=_vp(x0), x1=_vp(x1),
                          mask=_vp(mask), w=_vp(w), bias=_fp(bias), bias_per_sample=1 if bias_per_sample else 0,
                          relu=1 if relu else 0, y0=_vp(y0), y1=_vp(y1), head=_fp(head), prob=_fp(prob),
                          pool=_vp(pool))
        _lib.call("cm21_unet_conv3_fwd", ctypes.byref(c), self._st())

    def _wgrad(self, x0, g, mask, name, cin, cout, D, x1=None, c0=None):
        w = _lib.UnetWgrad(B=self.B, D=D, cin=cin, cout=cout, c0=c0 or cin, x0=_vp(x0), x1=_vp(x1), g=_vp(g),
                           mask=_vp(mask), dw=_fp(self.g[name + ".w"]), db=_fp(self.g[name + ".b"]),
                           work=_vp(self.work), work_bytes=self.work.numel())
        _lib.call("cm21_unet_conv3_wgrad", ctypes.byref(w), self._st())

    # -------------------------------------------------------------- passes
    def forward(self, delta, cond, keep=True):
        """delta: (B,D,D,D) fp32 CUDA tensor; cond: (B,2) normalised
        conditioning.  Returns the probability tensor (B,D,D,D) (a buffer).
        keep=False (inference) fuses the head into the last convolution,
        whose activations are then not stored, so backward() needs a
        forward with keep=True; both give the same probabilities."""
        if tuple(delta.shape) != (self.B, self.D, self.D, self.D) or delta.dtype != torch.float32:
            raise ValueError(f"delta must be float32 {(self.B, self.D, self.D, self.D)}")
        if not delta.is_contiguous():
            raise ValueError("delta must be contiguous")
        self.delta = delta
        self.cond.copy_(torch.as_tensor(cond, dtype=torch.float32).reshape(self.B, 2))
        a, p, D = self.a, self.p, self.D
        h, q = D // 2, D // 4
        st = self._st()
        _lib.call("cm21_unet_in_fwd", _fp(delta), _fp(self.cond), _fp(p["in.w"]), _fp(p["in.b"]), _fp(self.beff),
                  _vp(a["e1a"]), self.B, D, ctypes.c_float(DELTA_SCALE), st)
        # each encoder level's last conv also writes the pooled box
        self._conv(a["e1a"], self.wb["e1"], a["e1"], 16, 16, D, bias=p["e1.b"], pool=a["p1"])
        self._conv(a["p1"], self.wb["e2a"], a["e2a"], 16, 32, h, bias=p["e2a.b"])
        self._conv(a["e2a"], self.wb["e2b"], a["e2"], 32, 32, h, bias=p["e2b.b"], pool=a["p2"])
        self._conv(a["p2"], self.wb["b1"], a["b1"], 32, 64, q, bias=p["b1.b"])
        self._conv(a["b1"], self.wb["b2"], a["b2"], 64, 64, q, bias=p["b2.b"])
        _lib.call("cm21_unet_up_fwd", _vp(a["b2"]), _vp(self.wb["up2"]), _fp(p["up2.b"]), _vp(a["u2"]), self.B, q, 64,
                  32, st)
        self._conv(a["e2"], self.wb["d2a"], a["d2a"], 64, 32, h, x1=a["u2"], c0=32, bias=p["d2a.b"])
        self._conv(a["d2a"], self.wb["d2b"], a["d2"], 32, 32, h, bias=p["d2b.b"])
        _lib.call("cm21_unet_up_fwd", _vp(a["d2"]), _vp(self.wb["up1"]), _fp(p["up1.b"]), _vp(a["u1"]), self.B, h, 32,
                  16, st)
        self._conv(a["e1"], self.wb["d1a"], a["d1a"], 32, 16, D, x1=a["u1"], c0=16, bias=p["d1a.b"])
        self._kept = bool(keep)
        if keep:
            self._conv(a["d1a"], self.wb["d1b"], a["d1"], 16, 16, D, bias=p["d1b.b"])
            _lib.call("cm21_unet_head", _vp(a["d1"]), _fp(p["head.w"]), _fp(p["head.b"]), None, _fp(self.prob), None,
                      None, None, self.B, D, None, 0, st)
        else:
            # head.w (16) and head.b (1) are adjacent in the flat parameter buffer
            self._conv(a["d1a"], self.wb["d1b"], None, 16, 16, D, bias=p["d1b.b"], head=p["head.w"], prob=self.prob)
        return self.prob

    def backward(self, target):
        """Loss and gradients of the last forward against target (B,D,D,D)
        fp32; fills self.grad.  Returns the loss as a device tensor."""
        if not getattr(self, "_kept", False):
            raise RuntimeError("backward() needs the activations of a forward(..., keep=True)")
        a, ga, p, D = self.a, self.ga, self.p, self.D
        h, q = D // 2, D // 4
        st = self._st()
        head_dwb = self.g["head.w"]  # head.w (16) and head.b (1) are adjacent in the flat buffer
        _lib.call("cm21_unet_head", _vp(a["d1"]), _fp(p["head.w"]), _fp(p["head.b"]), _fp(target), _fp(self.prob),
                  _vp(ga["d1"]), _fp(head_dwb), ctypes.cast(ctypes.c_void_p(self.loss.data_ptr()),
                                                            ctypes.POINTER(ctypes.c_double)),
                  self.B, D, _vp(self.work), self.work.numel(), st)
        # level 1 decoder
        self._wgrad(a["d1a"], ga["d1"], a["d1"], "d1b", 16, 16, D)
        self._conv(ga["d1"], self.wt["d1b"], ga["d1a"], 16, 16, D, mask=a["d1"], relu=False)
        self._wgrad(a["e1"], ga["d1a"], a["d1a"], "d1a", 32, 16, D, x1=a["u1"], c0=16)
        self._conv(ga["d1a"], self.wt["d1a"], ga["e1"], 16, 32, D, mask=a["d1a"], relu=False, y1=ga["u1"], o0=16)
        _lib.call("cm21_unet_up_bwd", _vp(a["d2"]), _vp(ga["u1"]), _vp(self.wt["up1"]), _vp(ga["d2"]), _fp(self.g["up1.w"]),
                  _fp(self.g["up1.b"]), self.B, h, 32, 16, _vp(self.work), self.work.numel(), st)
        # level 2 decoder
        self._wgrad(a["d2a"], ga["d2"], a["d2"], "d2b", 32, 32, h)
        self._conv(ga["d2"], self.wt["d2b"], ga["d2a"], 32, 32, h, mask=a["d2"], relu=False)
        self._wgrad(a["e2"], ga["d2a"], a["d2a"], "d2a", 64, 32, h, x1=a["u2"], c0=32)
        self._conv(ga["d2a"], self.wt["d2a"], ga["e2"], 32, 64, h, mask=a["d2a"], relu=False, y1=ga["u2"], o0=32)
        _lib.call("cm21_unet_up_bwd", _vp(a["b2"]), _vp(ga["u2"]), _vp(self.wt["up2"]), _vp(ga["b2"]), _fp(self.g["up2.w"]),
                  _fp(self.g["up2.b"]), self.B, q, 64, 32, _vp(self.work), self.work.numel(), st)
        # bottleneck
        self._wgrad(a["b1"], ga["b2"], a["b2"], "b2", 64, 64, q)
        self._conv(ga["b2"], self.wt["b2"], ga["b1"], 64, 64, q, mask=a["b2"], relu=False)
        self._wgrad(a["p2"], ga["b1"], a["b1"], "b1", 32, 64, q)
        self._conv(ga["b1"], self.wt["b1"], ga["p2"], 64, 32, q, mask=a["b1"], relu=False)
        _lib.call("cm21_unet_pool_bwd", _vp(ga["p2"]), _vp(ga["e2"]), self.B, h, 32, 1, st)
        # level 2 encoder
        self._wgrad(a["e2a"], ga["e2"], a["e2"], "e2b", 32, 32, h)
        self._conv(ga["e2"], self.wt["e2b"], ga["e2a"], 32, 32, h, mask=a["e2"], relu=False)
        self._wgrad(a["p1"], ga["e2a"], a["e2a"], "e2a", 16, 32, h)
        self._conv(ga["e2a"], self.wt["e2a"], ga["p1"], 32, 16, h, mask=a["e2a"], relu=False)
        _lib.call("cm21_unet_pool_bwd", _vp(ga["p1"]), _vp(ga["e1"]), self.B, D, 16, 1, st)
        # level 1 encoder
        self._wgrad(a["e1a"], ga["e1"], a["e1"], "e1", 16, 16, D)
        self._conv(ga["e1"], self.wt["e1"], ga["e1a"], 16, 16, D, mask=a["e1"], relu=False)
        _lib.call("cm21_unet_in_wgrad", _fp(self.delta), _fp(self.cond), _vp(ga["e1a"]), _vp(a["e1a"]),
                  _fp(self.g["in.w"]), _fp(self.g["in.b"]), self.B, D, ctypes.c_float(DELTA_SCALE),
                  _vp(self.work), self.work.numel(), st)
        return self.loss

    def step(self, lr, beta1=0.9, beta2=0.999, eps=1e-8, weight_decay=1e-4):
        self.step_count += 1
        _lib.call("cm21_adamw", _fp(self.flat), _fp(self.grad), _fp(self.m), _fp(self.v), self.nparams,
                  ctypes.c_float(lr), ctypes.c_float(beta1), ctypes.c_float(beta2), ctypes.c_float(eps),
                  ctypes.c_float(weight_decay), self.step_count, self._st())
        self.pack()


# ------------------------------------------------------------------ data

def surrogate_specs(seeds, seed=7, D=128, box_len=200.0, deltac=4.0):
    """Per-box parameters of a surrogate data set: zeta and sigma_min^2 drawn
    uniformly (NumPy PCG64(seed)) for the given GRF seeds; a fixed barrier."""
    rng = np.random.Generator(np.random.PCG64(int(seed)))
    out = []
    for s in seeds:
        out.append({"seed": int(s), "D": int(D), "box_len": float(box_len), "deltac": float(deltac),
                    "zeta": float(rng.uniform(*ZETA_RANGE)), "sigma_min2": float(rng.uniform(*SMIN2_RANGE))})
    return out


class BoxSet:
    """Density boxes and the plain method's ionized fields, resident in HBM."""

    def __init__(self, specs, chunk=64, stream=None):
        g = importlib.import_module(__package__ + ".gpu")
        sig = importlib.import_module(__package__ + ".cosmo_files.sigmas")
        bm = importlib.import_module(__package__ + ".Programs.find_HII_bubbles_batched")
        self.specs = list(specs)
        D, L = self.specs[0]["D"], self.specs[0]["box_len"]
        if any(s["D"] != D or s["box_len"] != L for s in self.specs):
            raise ValueError("one (D, box_len) per data set")
        self.D, self.L = D, L
        n = len(self.specs)
        dev = torch.device("cuda")
        self.delta = torch.empty((n, D, D, D), dtype=torch.float32, device=dev)
        self.target = torch.empty((n, D, D, D), dtype=torch.float32, device=dev)
        self.cond = torch.tensor([normalise_cond(s["zeta"], s["sigma_min2"]) for s in self.specs],
                                 dtype=torch.float32)
        amp = sig.power_law_spectrum(D, L).amp
        for i, s in enumerate(self.specs):
            self.delta[i].copy_(g.c2r(g.grf_init(D, L, amp, s["seed"])))
        for c0 in range(0, n, chunk):
            sub = self.specs[c0:c0 + chunk]
            bb = bm.BatchedBubbles(sub, fil=1, LE=True, graph=False, tiles=None, stream=stream)
            for j in range(len(sub)):
                bb.set_delta(j, self.delta[c0 + j])
            bb.run()
            for j in range(len(sub)):
                self.target[c0 + j].copy_(bb.ionized(j))
            del bb

    def __len__(self):
        return len(self.specs)

    def x_HI(self):
        return 1.0 - self.target.view(len(self), -1).double().mean(1).cpu().numpy()


# ------------------------------------------------------------------ metrics

def box_metrics(field_sur, field_plain, delta, box_len, nbins=40):
    """|dx_HI|, cell agreement (ionized = field >= 0.5) and the P21(k) relative
    error (max, and mean, over the bins with power above 1e-3 of the peak) of
    one box."""
    dt = importlib.import_module(__package__ + ".Programs.delta_T")
    D = field_plain.shape[0]
    xs = 1.0 - float(field_sur.double().mean())
    xp = 1.0 - float(field_plain.double().mean())
    agree = float(((field_sur >= 0.5) == (field_plain >= 0.5)).double().mean())
    edges = dt.log_edges(box_len, D, nbins)
    ps = []
    for f in (field_sur, field_plain):
        dT, s = dt.brightness_temperature((1.0 - f).contiguous(), delta, 8.0)
        _, _, pk, _ = dt.power_spectrum(dT, box_len, edges, field_sum=s)
        ps.append(pk)
    ok = ps[1] > 1e-3 * max(ps[1].max(), 1e-300)
    rel = np.abs(ps[0][ok] - ps[1][ok]) / ps[1][ok]
    p21 = float(np.max(rel)) if ok.any() else 0.0
    p21_mean = float(np.mean(rel)) if ok.any() else 0.0
    return {"dx_HI": abs(xs - xp), "x_HI_plain": xp, "x_HI_sur": xs, "agreement": agree, "p21_err": p21,
            "p21_err_mean_bins": p21_mean}


def evaluate(model, boxes, idx=None):
    """Metrics of the model against the plain fields of a BoxSet."""
    idx = list(range(len(boxes))) if idx is None else list(idx)
    out = []
    for i0 in range(0, len(idx), model.B):
        sel = idx[i0:i0 + model.B]
        if len(sel) < model.B:
            sel = sel + [sel[-1]] * (model.B - len(sel))
        delta = boxes.delta[sel].contiguous()
        prob = model.forward(delta, boxes.cond[sel], keep=False)
        for j, i in enumerate(idx[i0:i0 + model.B]):
            out.append(box_metrics(prob[j], boxes.target[i], boxes.delta[i], boxes.L))
    return out


def summarise(metrics, tol_xhi=0.01, tol_p21=0.10):
    dx = np.array([m["dx_HI"] for m in metrics])
    ag = np.array([m["agreement"] for m in metrics])
    p21 = np.array([m["p21_err"] for m in metrics])
    p21m = np.array([m["p21_err_mean_bins"] for m in metrics])
    return {"boxes": len(metrics), "dx_HI_mean": float(dx.mean()), "dx_HI_max": float(dx.max()),
            "frac_dx_HI_within_tol": float(np.mean(dx < tol_xhi)),
            "agreement_mean": float(ag.mean()), "p21_err_median": float(np.median(p21)),
            "p21_err_mean_bins_median": float(np.median(p21m)),
            "frac_within_tol": float(np.mean((dx < tol_xhi) & (p21 < tol_p21))),
            "frac_within_tol_mean_bins": float(np.mean((dx < tol_xhi) & (p21m < tol_p21)))}


# ------------------------------------------------------------------ training

def _symmetry(delta, target, rng):
    """One random symmetry of the periodic cube applied to a batch of
    (B, D, D, D) boxes and their targets: an axis permutation, reflections
    and a periodic shift (the excursion-set field is equivariant under all
    of them: the filters depend on |k| only)."""
    perm = [int(a) + 1 for a in rng.permutation(3)]
    delta = delta.permute(0, *perm)
    target = target.permute(0, *perm)
    flips = [a + 1 for a in range(3) if rng.random() < 0.5]
    if flips:
        delta, target = delta.flip(flips), target.flip(flips)
    D = delta.shape[1]
    shift = tuple(int(v) for v in rng.integers(0, D, size=3))
    return delta.roll(shift, dims=(1, 2, 3)), target.roll(shift, dims=(1, 2, 3))


def train(model, boxes, steps, lr=2e-3, lr_min=1e-4, seed=0, log_every=50, log=None, time_budget=None,
          augment=False, rng=None, start=0, checkpoint=None, checkpoint_every=0):
    """AdamW with a cosine learning-rate schedule over random batches of the
    BoxSet (order from NumPy PCG64(seed)), optionally with a random symmetry
    of the periodic cube per batch.  With time_budget (seconds) the number of
    steps is cut to what fits, measured on the first 10 steps.  A resumed run
    passes the generator load_checkpoint returned and the step to start from;
    with checkpoint (a path) a checkpoint is written every checkpoint_every
    steps and at the end.  Returns the logged (step, loss) pairs."""
    import time
    rng = rng if rng is not None else np.random.Generator(np.random.PCG64(int(seed)))
    losses = []
    n = len(boxes)
    t0 = None
    it = int(start)
    while it < int(steps):
        if time_budget and it == start + 10:
            torch.cuda.synchronize()
            per = (time.perf_counter() - t0) / 10
            steps = min(int(steps), start + 10 + int((time_budget - 10 * per) / per))
            if log:
                log(f"{per * 1e3:.1f} ms/step: training for {steps} steps")
        if it == start:
            torch.cuda.synchronize()
            t0 = time.perf_counter()
        sel = rng.choice(n, size=model.B, replace=False)
        delta, target = boxes.delta[sel], boxes.target[sel]
        if augment:
            delta, target = _symmetry(delta, target, rng)
        model.forward(delta.contiguous(), boxes.cond[sel])
        loss = model.backward(target.contiguous())
        cur = lr_min + 0.5 * (lr - lr_min) * (1.0 + math.cos(math.pi * it / max(int(steps) - 1, 1)))
        model.step(cur)
        if log_every and (it % log_every == 0 or it == steps - 1):
            lv = float(loss)
            losses.append((it, lv))
            if log:
                log(f"step {it} loss {lv:.5f} lr {cur:.2e}")
        it += 1
        if checkpoint and checkpoint_every and it % checkpoint_every == 0:
            model.save_checkpoint(checkpoint, rng=rng, step=it, steps=int(steps))
    if checkpoint:
        model.save_checkpoint(checkpoint, rng=rng, step=it, steps=int(steps))
    return losses


def calibrate_logit_scale(model, boxes, scales=(1.0, 1.25, 1.5, 1.75, 2.0, 2.5, 3.0), idx=None, log=None):
    """Temperature scaling of the output: multiply the head's weights and bias
    by the scale a (p = sigmoid(a * logit)) that minimises the median P21(k)
    error on the given (training) boxes, among those that keep the mean
    |dx_HI| within 10% of the unscaled one.  Applies the chosen scale to the
    model and returns (scale, {scale: summary})."""
    w0 = model.p["head.w"].clone()
    b0 = model.p["head.b"].clone()
    table = {}
    for a in scales:
        model.p["head.w"].copy_(w0 * a)
        model.p["head.b"].copy_(b0 * a)
        table[a] = summarise(evaluate(model, boxes, idx))
        if log:
            log(f"logit scale {a}: {table[a]}")
    base = table[scales[0]]["dx_HI_mean"]
    ok = [a for a in scales if table[a]["dx_HI_mean"] <= 1.1 * base + 1e-4]
    best = min(ok, key=lambda a: table[a]["p21_err_median"])
    model.p["head.w"].copy_(w0 * best)
    model.p["head.b"].copy_(b0 * best)
    return best, table


# ------------------------------------------------------------------ serving with fallback

class SurrogatePipeline:
    """Ionized fields for (delta, zeta, sigma_min^2) requests from the
    network, with the plain method as fallback and judge: every
    `validate_every`-th box is also computed with the plain method; if it
    misses the tolerance the whole request is recomputed with the plain
    method and counted in `fallbacks`.  After `max_failures` failures the
    network is bypassed for all later requests."""

    def __init__(self, model, box_len=200.0, deltac=4.0, validate_every=8, tol_xhi=0.01, tol_p21=0.10,
                 max_failures=3):
        self.model = model
        self.box_len, self.deltac = float(box_len), float(deltac)
        self.validate_every = int(validate_every)
        self.tol_xhi, self.tol_p21 = float(tol_xhi), float(tol_p21)
        self.max_failures = int(max_failures)
        self.seen = self.validated = self.fallbacks = self.failures = 0
        self.disabled = False

    def _plain(self, delta, zeta, smin2):
        bm = importlib.import_module(__package__ + ".Programs.find_HII_bubbles_batched")
        D = delta.shape[-1]
        specs = [{"D": D, "box_len": self.box_len, "deltac": self.deltac, "zeta": float(z), "sigma_min2": float(s)}
                 for z, s in zip(zeta, smin2)]
        bb = bm.BatchedBubbles(specs, fil=1, LE=True, graph=False, tiles=None)
        for j in range(len(specs)):
            bb.set_delta(j, delta[j])
        bb.run()
        return torch.stack([bb.ionized(j).clone() for j in range(len(specs))])

    def __call__(self, delta, zeta, smin2):
        """delta (B,D,D,D) fp32 CUDA; zeta, smin2 sequences of B.  Returns
        (fields (B,D,D,D), 'surrogate' | 'plain')."""
        B = self.model.B
        if delta.shape[0] != B:
            raise ValueError(f"requests come in batches of {B}")
        if self.disabled:
            self.fallbacks += 1
            return self._plain(delta, zeta, smin2), "plain"
        cond = [normalise_cond(z, s) for z, s in zip(zeta, smin2)]
        prob = self.model.forward(delta.contiguous(), cond, keep=False).clone()
        check = [j for j in range(B) if (self.seen + j) % self.validate_every == 0]
        self.seen += B
        if check:
            ref = self._plain(delta[check], [zeta[j] for j in check], [smin2[j] for j in check])
            self.validated += len(check)
            for k, j in enumerate(check):
                m = box_metrics(prob[j], ref[k], delta[j], self.box_len)
                if m["dx_HI"] > self.tol_xhi or m["p21_err"] > self.tol_p21:
                    self.failures += 1
                    self.fallbacks += 1
                    if self.failures >= self.max_failures:
                        self.disabled = True
                    return self._plain(delta, zeta, smin2), "plain"
        return prob, "surrogate"


def _seed_range(text):
    a, b = (int(v) for v in text.split(":"))
    return range(a, b)


def main(argv=None):
    """Train the surrogate on plain-method boxes and evaluate it on held-out
    seeds: python -m 21cmpy_amd.surrogate --steps 4000 --out weights.npz"""
    import argparse
    import time
    ap = argparse.ArgumentParser(description="train / evaluate the U-Net surrogate")
    ap.add_argument("--D", type=int, default=128)
    ap.add_argument("--box-len", type=float, default=200.0)
    ap.add_argument("--deltac", type=float, default=4.0)
    ap.add_argument("--train-seeds", default="1000:3000")
    ap.add_argument("--eval-seeds", default="3000:3200")
    ap.add_argument("--batch", type=int, default=4)
    ap.add_argument("--steps", type=int, default=4000)
    ap.add_argument("--lr", type=float, default=2e-3)
    ap.add_argument("--lr-min", type=float, default=1e-4)
    ap.add_argument("--augment", action="store_true", help="random cube symmetry per batch")
    ap.add_argument("--time-budget", type=float, default=0.0, help="seconds of training at most")
    ap.add_argument("--init", default=None, help="start from these weights")
    ap.add_argument("--out", default=None, help="weights file to write")
    ap.add_argument("--eval-only", action="store_true")
    ap.add_argument("--calibrate", type=int, default=0,
                    help="fit the output temperature on this many training boxes after training")
    args = ap.parse_args(argv)

    def log(msg):
        print(msg, flush=True)

    t = time.perf_counter()
    params = UNet3D.load_file(args.init)[0] if args.init else None
    model = UNet3D(D=args.D, batch=args.batch, params=params)
    report = {"D": args.D, "box_len": args.box_len, "deltac": args.deltac, "batch": args.batch}
    if not args.eval_only:
        train_set = BoxSet(surrogate_specs(_seed_range(args.train_seeds), seed=7, D=args.D, box_len=args.box_len,
                                           deltac=args.deltac))
        xh = train_set.x_HI()
        log(f"training boxes: {len(train_set)} in {time.perf_counter() - t:.1f} s; x_HI quartiles "
            f"{np.round(np.quantile(xh, [0, 0.25, 0.5, 0.75, 1]), 3).tolist()}")
        t = time.perf_counter()
        losses = train(model, train_set, args.steps, lr=args.lr, lr_min=args.lr_min, log_every=50, log=log,
                       time_budget=args.time_budget or None, augment=args.augment)
        torch.cuda.synchronize()
        report["train"] = {"boxes": len(train_set), "steps": losses[-1][0] + 1, "seconds": time.perf_counter() - t,
                           "loss_first": losses[0][1], "loss_last": losses[-1][1]}
        if args.calibrate:
            a, table = calibrate_logit_scale(model, train_set, idx=range(min(args.calibrate, len(train_set))), log=log)
            report["logit_scale"] = a
            log(f"output temperature: logit scale {a}")
        del train_set
        torch.cuda.empty_cache()
        if args.out:
            model.save(args.out, **{k: v for k, v in report.items() if k != "train"})
            log(f"weights written to {args.out}")
    elif args.calibrate:
        # calibration on the first training seeds only (the same draws as training)
        seeds = list(_seed_range(args.train_seeds))
        specs = surrogate_specs(seeds, seed=7, D=args.D, box_len=args.box_len, deltac=args.deltac)
        cal_set = BoxSet(specs[:args.calibrate])
        a, _ = calibrate_logit_scale(model, cal_set, log=log)
        report["logit_scale"] = a
        log(f"output temperature: logit scale {a}")
        del cal_set
        if args.out:
            model.save(args.out, **report)
            log(f"weights written to {args.out}")
    eval_set = BoxSet(surrogate_specs(_seed_range(args.eval_seeds), seed=8, D=args.D, box_len=args.box_len,
                                      deltac=args.deltac))
    report["eval"] = summarise(evaluate(model, eval_set))
    log(json.dumps(report))
    return report


if __name__ == "__main__":
    main()
