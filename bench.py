"""Benchmark of the find_HII_bubbles hot path on one MI355X.

Workload (the full plain pipeline at HII_DIM=256): a DIM=512, BOX_LEN=300 Mpc
Gaussian random field (P(k) ~ k^-2.5 exp(-(k/k_nyq)^2), per-cell sigma 0.6,
seed 0) is smoothed and nearest-picked to 256^3, moved by the Zel'dovich
displacement at growth 0.12 (D_init = 0.12*9/301) with a cloud-in-cell
deposit, and the resulting nonlinear 256^3 box goes through find_HII_bubbles:
40 radii, zeta = 30, k-space top-hat (HII_FILTER = 1), box-mean
normalisation, sigma_min^2 = 4.0, barrier 1.686/0.12.

A step is one find_HII_bubbles pass over that box, with the box already in
HBM.  The JSON line reports boxes/s and, for the dominant kernel (the
pipelined radius launch), its roofline by the unit byte model (algorithmic
bytes per launch over its HIP-event-timed average duration; the fused loop
moves far fewer bytes than the model counts, so the fraction can exceed 1),
the calibrated HBM traffic per launch and the VALU issue rate (both from
committed rocprofv3 counter summaries under profiles/).  Beside the headline:
a partly ionized variant of the same box, the 512^3 and 64^3 (literal
parameters) boxes, the 64-box batched path (literal and partly ionized), the
U-Net surrogate pipeline with its validation, and a CPU baseline: the NumPy
FP32 oracle of the same algorithm on a bounded sample (one core and all
cores).  --no-extra / --no-batched / --no-surrogate / --no-cpu-baseline skip
those parts.
"""
import argparse
import importlib
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0  # MI355X HBM3E, published
# VALU issue peak of MI355X for wave64 FP32 instructions: 256 CUs x 4 SIMDs x
# 2.4 GHz, four cycles per instruction (packed FP32 is not used)
VALU_PEAK_G = 256 * 4 * 2.4 / 4


def alg_bytes(D, norm=1):
    """Algorithmic bytes per launch of each stage, from the unit model
    B_alg = 4N + 8Nc + n_R (24 Nc + (16 + 4 norm) N): filter 16Nc,
    c2r 8Nc + 4N, f_coll box sum 4N, threshold 12N.  A fused launch carries the
    model bytes of the stages it replaces: one pipelined radius launch holds
    one radius' worth of every stage (the filter in its columns role, the c2r
    across its columns, rows and lines roles, the box sum and the threshold),
    24Nc + (16 + 4 norm) N; the final paint carries the last threshold, 12N."""
    N = D ** 3
    Nc = D * D * (D // 2 + 1)
    return {
        "r2c": 4 * N + 8 * Nc,
        "filter": 16 * Nc,
        "c2r": 8 * Nc + 4 * N,
        "fcoll_sum": 4 * N,
        "threshold": 12 * N,
        "fused_radius": 24 * Nc + (16 + 4 * norm) * N,
        "fused_paint": 12 * N,
    }


def b_alg(D, n_r, norm):
    N = D ** 3
    Nc = D * D * (D // 2 + 1)
    return 4 * N + 8 * Nc + n_r * (24 * Nc + (16 + 4 * norm) * N)


def cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_loop(delta, L, radii, denoms, zeta, deltac, fc_mean, LE, workers, budget_s):
    """The oracle's FP32 radius loop (oracle/bubbles.py, the same windows,
    f_coll and threshold) with scipy.fft transforms on `workers` threads, on
    the host cores: the forward transform plus as many radii as fit the
    budget (all of them when budget_s is None).  Returns (seconds per box
    extrapolated to every radius, radii timed, ionized box)."""
    import scipy.fft as sfft
    from oracle import bubbles
    D = delta.shape[0]
    t0 = time.perf_counter()
    d = delta if LE else (delta * np.float32(deltac / 1.686)).astype(np.float32)
    spec = sfft.rfftn(d.astype(np.float64), workers=workers).astype(np.complex64)
    km = bubbles.kmag_half_f32(D, bubbles.delta_k_f32(L))
    t_fwd = time.perf_counter() - t0
    done, t_rad = 0, 0.0
    ion = np.zeros(delta.shape, np.float32)
    for r, (R, den) in enumerate(zip(radii, denoms)):
        t1 = time.perf_counter()
        w0, cut, w2 = bubbles.window_f32(km, R, 1, True)
        filt = bubbles.apply_window_f32(spec, w0, cut, w2)
        sm = sfft.irfftn(filt.astype(np.complex128), s=(D, D, D), axes=(0, 1, 2), workers=workers).astype(np.float32)
        fc = bubbles.fcoll_f32(sm, deltac, den)
        if not LE:
            fc = (fc * bubbles.box_mean_scale_f32(fc, fc_mean)).astype(np.float32)
        hit = fc.astype(np.float64) >= 1.0 / zeta
        if r == len(radii) - 1:
            todo = ion != 1.0
            ion[todo & hit] = 1.0
            part = todo & ~hit
            ion[part] = (fc[part].astype(np.float64) * zeta).astype(np.float32)
        else:
            ion[hit] = 1.0
        t_rad += time.perf_counter() - t1
        done += 1
        if budget_s is not None and t_fwd + t_rad > budget_s:
            break
    return t_fwd + t_rad / done * len(radii), done, ion


def cpu_baseline(delta, L, radii, denoms, zeta, deltac, fc_mean, budget_s=15.0, gpu_config1=None):
    """The same algorithm on the host cores (kind "port": the oracle's FP32
    restatement), timed on a bounded sample of the benchmark box with scipy.fft
    on every available core (the reported value) and on one thread, plus
    configuration 1 (64^3, 37 radii, linear input) timed in full."""
    from oracle import grf, sigma
    ncores = len(os.sched_getaffinity(0))
    D = delta.shape[0]
    per_box_mt, done_mt, _ = cpu_loop(delta, L, radii, denoms, zeta, deltac, fc_mean, False, ncores, budget_s)
    per_box_1, done_1, _ = cpu_loop(delta, L, radii, denoms, zeta, deltac, fc_mean, False, 1, budget_s)
    # configuration 1 in full: D=64, L=100, seed 0, sigma_min^2 = 4, zeta = 30, deltac = 1.686, linear input
    d1 = grf.c2r(grf.grf_half_spectrum(64, 100.0, seed=0), 64).astype(np.float32)
    r1, _ = sigma.radius_schedule(100.0, 64)
    den1 = sigma.denominators(r1, 4.0, lambda R: sigma.sig0_discrete(R, 64, 100.0))
    t0 = time.perf_counter()
    _, _, ion1 = cpu_loop(d1, 100.0, r1, den1, 30.0, 1.686, 0.0, True, 1, None)
    t_c1 = time.perf_counter() - t0
    out = {"value": 1.0 / per_box_mt, "unit": "boxes/s", "cores": ncores, "kind": "port",
           "cpu_model": cpu_model(), "threads": {"scipy.fft workers": ncores, "numpy elementwise": 1},
           "sample": f"oracle FP32 radius loop (oracle/bubbles.py windows, f_coll, threshold) with scipy.fft on "
                     f"{ncores} cores on the same {D}^3 benchmark box: forward transform + {done_mt} of "
                     f"{len(radii)} radii timed (budget {budget_s:.0f} s), extrapolated to one box",
           "single_thread": {"value": 1.0 / per_box_1, "unit": "boxes/s", "cores": 1,
                             "sample": f"{done_1} of {len(radii)} radii, scipy.fft workers=1, extrapolated"},
           "config1_full": {"value": 1.0 / t_c1, "unit": "boxes/s", "cores": 1, "seconds": t_c1,
                            "n_radii": len(r1), "x_HI": float(1.0 - ion1.mean()),
                            "workload": "configuration 1: 64^3 GRF seed 0, BOX_LEN 100, 37 radii, sigma_min^2 4, "
                                        "zeta 30, linear input, every radius timed, one thread"}}
    if gpu_config1:
        out["config1_full"]["gpu_boxes_s"] = gpu_config1
    return out


def loop_only(args, lib, torch):
    """The radius loop alone on a GRF box (D=256, L=300) scaled by the growth
    0.12 as nonlinear input: a short program for counter collection."""
    g = importlib.import_module("21cmpy_amd.gpu")
    fb = importlib.import_module("21cmpy_amd.Programs.find_HII_bubbles")
    sig = importlib.import_module("21cmpy_amd.cosmo_files.sigmas")
    D, L, growth = 256, 300.0, 0.12
    sp = sig.PowerLawSpectrum(D, L)
    delta = (g.c2r(g.grf_init(D, L, sp.amp, 0)) * growth).contiguous()
    kw = dict(LE=False, quiet=True, deltac=1.686 / growth, sigma_min2=4.0, zeta=30.0, sig0=sp.sig0_discrete,
              fused=not args.generic)
    for _ in range(max(args.steps, 1)):
        fb.conv_bubbles(delta, {"BoxSize": L, "HIIdim": D}, **kw)
    torch.cuda.synchronize()
    print(json.dumps({"mode": "loop", "D": D, "steps": args.steps}))


def partly_ionized_specs(specs, seed=1):
    """The manifest's boxes with sigma_min^2 chosen per box so that the mean
    collapsed fraction is zeta erfc(deltac/sqrt(2 sigma_min^2)) = c, c uniform
    in [0.3, 0.7] from NumPy PCG64(seed): on these linear sigma-0.6 boxes the
    neutral fraction comes out near 1 - c."""
    from scipy.special import erfcinv
    bm = importlib.import_module("21cmpy_amd.Programs.find_HII_bubbles_batched")
    rng = np.random.Generator(np.random.PCG64(int(seed)))
    out = []
    for s in specs:
        s = dict(s)
        dc = 1.686 / bm.linear_growth(s["z"])
        s["sigma_min2"] = (dc / float(erfcinv(rng.uniform(0.3, 0.7) / s["zeta"]))) ** 2 / 2.0
        out.append(s)
    return out


def batched_bench(args, torch, partial=False):
    """The 64-box batch (mixed HII_DIM 96/128/160, per-box zeta, z, sigma_min^2
    from the committed manifest; linear z=0 GRF boxes, cell 1.5625 Mpc, 37
    radii, HII_FILTER 1) through cm21_find_bubbles_batched as one graph; with
    partial, the same boxes with sigma_min^2 chosen so that each is partly
    ionized (partly_ionized_specs).  Returns the result object for the JSON
    line."""
    bm = importlib.import_module("21cmpy_amd.Programs.find_HII_bubbles_batched")
    g = importlib.import_module("21cmpy_amd.gpu")
    sig = importlib.import_module("21cmpy_amd.cosmo_files.sigmas")
    lib = importlib.import_module("21cmpy_amd._lib")
    specs = bm.load_manifest()["boxes"]
    if partial:
        specs = partly_ionized_specs(specs)
    stream = torch.cuda.Stream()
    with torch.cuda.stream(stream):
        bb = bm.BatchedBubbles(specs, fil=1, LE=True, graph=not args.no_graph, stream=stream)
        amps = {}
        for i, s in enumerate(specs):
            key = (s["D"], s["box_len"])
            if key not in amps:
                amps[key] = sig.PowerLawSpectrum(*key).amp
            bb.set_delta(i, g.c2r(g.grf_init(s["D"], s["box_len"], amps[key], s["seed"])))
        tuned = None if (args.no_tune or partial) else bb.autotune(reps=3, save_path=args.save_tiles)
    torch.cuda.synchronize()
    for _ in range(max(args.warmup, 1)):
        bb.run()
    torch.cuda.synchronize()
    lib.profile_collect()
    lib.profile_enable(True)
    bb.run()
    torch.cuda.synchronize()
    lib.profile_enable(False)
    stages = {k: {"ms": ms, "launches": n} for k, (ms, n) in lib.profile_collect().items()}
    steps = max(args.steps, 1)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        bb.run()
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0
    nb = len(specs)
    value = nb * steps / elapsed
    fb = importlib.import_module("21cmpy_amd.Programs.find_HII_bubbles")
    balg = sum(b_alg(s["D"], len(fb.radius_schedule(s["box_len"], s["D"])[0]), 0) for s in specs)
    xhi = bb.x_HI()
    traffic = read_traffic("batched_find_bubbles_mix64_traffic.txt", "batch_bytes")
    return {
        "value": value, "unit": "boxes/s", "boxes_per_batch": nb, "steps": steps,
        "ms_per_batch": elapsed * 1e3 / steps,
        "workload": "64 linear GRF boxes, HII_DIM in {96,128,160} (21/25/18), BOX_LEN = 1.5625 Mpc x HII_DIM, "
                    "37 radii, per-box zeta in [10,60], z in {7..10} (deltac = 1.686/D(z), D(z)=0.12*9/(1+z)), "
                    + ("sigma_min^2 with zeta <f_coll> in [0.3,0.7] (partly ionized)" if partial else
                       "sigma_min^2 in [2,6]") + ", HII_FILTER 1; one graph per batch",
        # the unit model credits every box with the full per-radius passes;
        # the batched loop prunes them and keeps the mask on chip, so this
        # ratio is a model-bytes rate, not a bandwidth
        "model_bytes_rate": {"B_alg_per_batch": balg, "model_bytes_per_s_over_hbm_peak":
                             balg * steps / elapsed / (HBM_PEAK_GBS * 1e9)},
        "hbm_traffic": ({"bytes_per_batch": traffic, "gbs": traffic * steps / elapsed / 1e9,
                         "frac": traffic * steps / elapsed / (HBM_PEAK_GBS * 1e9),
                         "source": "profiles/batched_find_bubbles_mix64_traffic.txt (literal parameters)"}
                        if traffic and not partial else None),
        "stages_one_batch_eager": stages,
        "tiles": ({k: v["chosen"] for k, v in tuned.items()} if tuned else
                  {k: {"columns": c, "lines": l} for k, (c, l) in bb.tiles().items()}),
        "x_HI_mean": float(np.mean(xhi)), "x_HI_min": float(np.min(xhi)),
    }


def read_traffic(name, key):
    """A calibrated HBM byte count from a committed rocprofv3 counter summary
    under profiles/ (None when absent)."""
    path = os.path.join(ROOT, "profiles", name)
    if not os.path.exists(path):
        return None
    for line in open(path):
        if line.startswith(key + ":"):
            return float(line.split(":", 1)[1].split()[0])
    return None


def unet_flops(D):
    """FLOPs of one surrogate forward pass at side D: 2 x MACs of the 3^3
    convolutions (matrix cores), the transposed convolutions, the input
    layer and the head."""
    h, q = D // 2, D // 4
    conv = [(D, 16, 16), (h, 16, 32), (h, 32, 32), (q, 32, 64), (q, 64, 64), (h, 64, 32), (h, 32, 32),
            (D, 32, 16), (D, 16, 16)]
    mfma = sum(2 * 27 * ci * co * s ** 3 for s, ci, co in conv)
    other = 2 * 27 * 16 * D ** 3 + 2 * 64 * 32 * h ** 3 + 2 * 32 * 16 * D ** 3 + 2 * 16 * D ** 3
    return mfma, other


def plain_rate(torch, delta, L, deltac, smin, zeta, steps, linear=True, fused=True, graph=True, warmup=3):
    """Boxes/s of the plain single-box loop (graph replay unless graph=False)
    on one resident box, radii and denominators computed beforehand; returns
    (boxes/s, ionized box)."""
    lib = importlib.import_module("21cmpy_amd._lib")
    fb = importlib.import_module("21cmpy_amd.Programs.find_HII_bubbles")
    sig = importlib.import_module("21cmpy_amd.cosmo_files.sigmas")
    D = delta.shape[0]
    radii, denoms, fc_mean = fb.bubble_setup(D, L, None, deltac, smin, sig.PowerLawSpectrum(D, L).sig0_discrete)
    prm = lib.BubbleParams(box_len=L, deltac=float(np.float32(deltac)), zeta=zeta,
                           growth_inv=float(np.float32(deltac / 1.686)), fc_mean=float(fc_mean),
                           linear_input=1 if linear else 0, filter_type=1, ref_fallthrough=1, update=0,
                           sphere_periodic=0, sphere_r_scale=1.0, fused=1 if fused else 0, graph=1 if graph else 0)
    plan = lib.plan(D, 1)
    nbytes = lib.query("cm21_find_bubbles_workspace_bytes", D, len(radii), 0)
    work = torch.empty(nbytes, dtype=torch.uint8, device="cuda")
    ion = torch.empty_like(delta)
    r_arr = np.ascontiguousarray(radii, np.float32)
    d_arr = np.ascontiguousarray(denoms, np.float32)
    stream = torch.cuda.Stream()
    stream.wait_stream(torch.cuda.current_stream())

    def step():
        lib.call("cm21_find_bubbles", plan.handle, lib.ptr(delta), lib.ptr(ion), r_arr.ctypes.data_as(lib.c_float_p),
                 d_arr.ctypes.data_as(lib.c_float_p), len(radii), prm, lib.ptr(work), nbytes,
                 lib.stream_handle(stream))

    torch.cuda.synchronize()
    for _ in range(warmup):
        step()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        step()
    torch.cuda.synchronize()
    rate = steps / (time.perf_counter() - t0)
    return rate, ion


def flag_compare(a, b):
    """Fraction of cells whose ionized flag differs and |x_HI(a) - x_HI(b)|."""
    fa, fb_ = (a == 1), (b == 1)
    flips = float((fa != fb_).double().mean())
    dx = abs(float(a.double().mean()) - float(b.double().mean()))
    return flips, dx


def extra_workloads(torch, upd, steps):
    """Throughput beside the headline: the headline box with a barrier that
    ionizes part of it (the threshold band and the mask writes do real work),
    checked fused against the rocFFT path; 512^3 boxes (literal and partly
    ionized parameters); configuration 1 (64^3) on the GPU."""
    g = importlib.import_module("21cmpy_amd.gpu")
    sig = importlib.import_module("21cmpy_amd.cosmo_files.sigmas")
    out = {}
    # headline box, nonlinear input, barrier 1.686/0.12, sigma_min^2 = 40: partly ionized
    L, growth, zeta, smin = 300.0, 0.12, 30.0, 40.0
    rate, ion = plain_rate(torch, upd, L, 1.686 / growth, smin, zeta, steps, linear=False)
    _, ion_g = plain_rate(torch, upd, L, 1.686 / growth, smin, zeta, 1, linear=False, fused=False, graph=False,
                          warmup=0)
    flips, dx = flag_compare(ion, ion_g)
    out["config2_partly_ionized"] = {
        "value": rate, "unit": "boxes/s", "x_HI": 1.0 - float(ion.double().mean()),
        "flag_flips_vs_rocfft_path": flips, "abs_dx_HI_vs_rocfft_path": dx,
        "workload": "the headline 256^3 nonlinear box, zeta 30, barrier 1.686/0.12, sigma_min^2 = 40 "
                    "(partly ionized), 40 radii, graph replay"}
    del ion, ion_g
    # 512^3, BOX_LEN 500, 41 radii, linear input, zeta 25 (configuration 4's box)
    D, L = 512, 500.0
    sp = sig.PowerLawSpectrum(D, L)
    box = g.c2r(g.grf_init(D, L, sp.amp, 0))
    for name, smin in (("plain_512_literal", 4.0), ("plain_512_partly_ionized", 0.6)):
        rate, ion = plain_rate(torch, box, L, 1.686, smin, 25.0, max(steps // 4, 3), linear=True, warmup=1)
        out[name] = {"value": rate, "unit": "boxes/s", "x_HI": 1.0 - float(ion.double().mean()),
                     "workload": f"512^3 linear GRF box (seed 0, BOX_LEN 500), 41 radii, zeta 25, "
                                 f"sigma_min^2 = {smin}, deltac 1.686, graph replay"}
        del ion
    del box
    torch.cuda.empty_cache()
    # configuration 1 on the GPU
    D, L = 64, 100.0
    sp = sig.PowerLawSpectrum(D, L)
    box = g.c2r(g.grf_init(D, L, sp.amp, 0))
    rate, ion = plain_rate(torch, box, L, 1.686, 4.0, 30.0, steps * 5, linear=True)
    out["config1"] = {"value": rate, "unit": "boxes/s", "x_HI": 1.0 - float(ion.double().mean()),
                      "workload": "configuration 1: 64^3 GRF seed 0, BOX_LEN 100, 37 radii, sigma_min^2 4, "
                                  "zeta 30, linear input, graph replay"}
    return out


def surrogate_bench(args, torch, plain_boxes_s=None):
    """U-Net surrogate at HII_DIM=128 (BOX_LEN=200 Mpc): inference boxes/s at
    batch 8, a training step's time, and accuracy against the plain method
    on held-out seeds (3000..3015; trained on 1000..2999 and 3200..7199)."""
    sur = importlib.import_module("21cmpy_amd.surrogate")
    D, L, B = 128, 200.0, 8
    trained = os.path.exists(sur.WEIGHTS)
    params, meta = sur.UNet3D.load_file(sur.WEIGHTS) if trained else (None, {})
    model = sur.UNet3D(D=D, batch=B, params=params)
    specs = sur.surrogate_specs(range(3000, 3016), seed=8, D=D, box_len=L)
    boxes = sur.BoxSet(specs)
    delta = boxes.delta[:B].contiguous()
    cond = boxes.cond[:B]
    for _ in range(2):
        model.forward(delta, cond, keep=False)
    steps = max(args.steps, 1)
    start, end = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    start.record()
    for _ in range(steps):
        model.forward(delta, cond, keep=False)
    end.record()
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0
    fwd_ms = start.elapsed_time(end) / steps
    value = B * steps / elapsed
    # a training step (forward, backward, AdamW) on a copy of the model
    tm = sur.UNet3D(D=D, batch=B, params=params)
    tgt = boxes.target[:B].contiguous()
    tm.forward(delta, cond)
    tm.backward(tgt)
    tm.step(1e-4)
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    for _ in range(3):
        tm.forward(delta, cond)
        tm.backward(tgt)
        tm.step(1e-4)
    torch.cuda.synchronize()
    train_ms = (time.perf_counter() - t1) * 1e3 / 3
    del tm
    summary = sur.summarise(sur.evaluate(model, boxes)) if trained else None
    # served through the pipeline: every 8th box validated against the plain
    # method (production tolerances |dx_HI| < 0.01, P21 < 10 %), a failed
    # validation answers the whole request with the plain method
    pipe = sur.SurrogatePipeline(model, box_len=L, validate_every=8)
    reqs = [(boxes.delta[i:i + B].contiguous(), [s_["zeta"] for s_ in specs[i:i + B]],
             [s_["sigma_min2"] for s_ in specs[i:i + B]]) for i in range(0, len(specs) - B + 1, B)]
    pipe(*reqs[0])
    pipe.seen = pipe.validated = pipe.fallbacks = pipe.failures = 0
    pipe.disabled = False
    torch.cuda.synchronize()
    t2 = time.perf_counter()
    nreq = 0
    for rep in range(max(steps // len(reqs), 1)):
        for r in reqs:
            pipe(*r)
            nreq += 1
    torch.cuda.synchronize()
    pipe_rate = nreq * B / (time.perf_counter() - t2)
    pipeline = {"value": pipe_rate, "unit": "boxes/s", "requests": nreq, "validate_every": 8,
                "validated": pipe.validated, "fallbacks": pipe.fallbacks, "disabled": pipe.disabled,
                "tolerances": {"dx_HI": pipe.tol_xhi, "p21": pipe.tol_p21}}
    # plain method on the same box size, one box per call through the C ABI
    # (graph replay), host set-up outside the timed region
    s0 = specs[0]
    plain128, _ = plain_rate(torch, boxes.delta[0].contiguous(), L, s0["deltac"], s0["sigma_min2"], s0["zeta"],
                             steps)
    mf, other = unet_flops(D)
    achieved = B * (mf + other) / (fwd_ms * 1e-3) / 1e12
    out = {
        "value": value, "unit": "boxes/s", "HII_DIM": D, "BOX_LEN": L, "batch": B, "steps": steps,
        "ms_per_forward_batch": fwd_ms, "ms_per_train_step_batch": train_ms,
        "trained": trained, "weights_meta": meta,
        "accuracy_vs_plain_heldout16": summary,
        "plain_same_HII_DIM_boxes_s": plain128,
        "pipeline_with_validation": pipeline,
        "roofline": {"bound": "mfma", "achieved": achieved, "peak": 2500.0, "unit": "TFLOP/s",
                     "frac": achieved / 2500.0, "flops_per_box": mf + other, "mfma_flops_per_box": mf},
        "workload": "3-level 16/32/64 3D U-Net, bf16 MFMA convs + fp32 accumulation, on linear GRF boxes "
                    "(zeta in [10,60], sigma_min^2 in [2,6], barrier 4.0)",
    }
    if plain_boxes_s:
        out["vs_plain_256"] = value / plain_boxes_s
        out["pipeline_with_validation"]["vs_plain_256"] = pipe_rate / plain_boxes_s
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-budget", type=float, default=15.0)
    ap.add_argument("--generic", action="store_true", help="rocFFT c2r + separate kernels instead of the fused path")
    ap.add_argument("--no-graph", action="store_true", help="launch eagerly instead of replaying a HIP graph")
    ap.add_argument("--no-batched", action="store_true", help="skip the 64-box batched measurement")
    ap.add_argument("--no-tune", action="store_true", help="batched: use the stored tile choice, no autotuning")
    ap.add_argument("--save-tiles", default=None, help="batched: merge the autotuned tiles into this JSON file")
    ap.add_argument("--no-surrogate", action="store_true", help="skip the U-Net surrogate measurement")
    ap.add_argument("--no-extra", action="store_true",
                    help="skip the partly ionized, 512^3 and configuration-1 throughput measurements")
    ap.add_argument("--mode", choices=["bench", "loop", "copy", "batched", "surrogate"], default="bench",
                    help="loop: only the radius loop on a GRF box (for counter runs); copy: a 512 MiB device "
                         "copy (to calibrate FETCH_SIZE/WRITE_SIZE)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {args.gpus} is not supported "
              "(independent boxes shard across GPUs with no exchange; see the design notes)", file=sys.stderr)
        sys.exit(2)

    import torch
    if not torch.cuda.is_available():
        print("bench.py needs a GPU", file=sys.stderr)
        sys.exit(1)
    if args.mode == "copy":
        n = 512 * 1024 * 1024 // 4
        src = torch.ones(n, device="cuda")
        dst = torch.empty_like(src)
        for _ in range(max(args.steps, 1)):
            dst.copy_(src)
        torch.cuda.synchronize()
        print(json.dumps({"mode": "copy", "bytes_per_launch_read": 4 * n, "bytes_per_launch_write": 4 * n}))
        return
    lib = importlib.import_module("21cmpy_amd._lib")
    if args.mode == "loop":
        return loop_only(args, lib, torch)
    if args.mode == "batched":
        print(json.dumps({"mode": "batched", "batched": batched_bench(args, torch)}))
        return
    if args.mode == "surrogate":
        print(json.dumps({"mode": "surrogate", "surrogate": surrogate_bench(args, torch)}))
        return
    init = importlib.import_module("21cmpy_amd.Programs.initialize")
    pert = importlib.import_module("21cmpy_amd.Programs.perturb_field")
    fb = importlib.import_module("21cmpy_amd.Programs.find_HII_bubbles")
    dt = importlib.import_module("21cmpy_amd.Programs.delta_T")
    sig = importlib.import_module("21cmpy_amd.cosmo_files.sigmas")

    DIM, D, L, seed = 512, 256, 300.0, 0
    growth, zeta, smin = 0.12, 30.0, 4.0
    growth_init = growth * 9.0 / 301.0
    deltac = 1.686 / growth

    def timed(fn):
        torch.cuda.synchronize()
        t = time.perf_counter()
        r = fn()
        torch.cuda.synchronize()
        return r, (time.perf_counter() - t) * 1e3

    # pipeline stages (once; the first call of each includes plan creation)
    ics, _ = timed(lambda: init.init(DIM=DIM, HII_DIM=D, BOX_LEN=L, seed=seed, pspec="powerlaw", save=False))
    ics, t_init = timed(lambda: init.init(DIM=DIM, HII_DIM=D, BOX_LEN=L, seed=seed, pspec="powerlaw", save=False))
    upd, _ = timed(lambda: pert.evolve_zeldovich(8.0, ics["deltax"], ics["v"], growth, growth_init, L, "cic", save=False))
    upd, t_pert = timed(lambda: pert.evolve_zeldovich(8.0, ics["deltax"], ics["v"], growth, growth_init, L, "cic", save=False))
    del ics
    torch.cuda.empty_cache()

    spectrum = sig.PowerLawSpectrum(D, L)
    radii, denoms, fc_mean = fb.bubble_setup(D, L, None, deltac, smin, spectrum.sig0_discrete)
    prm = lib.BubbleParams(box_len=L, deltac=float(np.float32(deltac)), zeta=zeta,
                           growth_inv=float(np.float32(deltac / 1.686)), fc_mean=float(fc_mean), linear_input=0,
                           filter_type=1, ref_fallthrough=1, update=0, sphere_periodic=0, sphere_r_scale=1.0,
                           fused=0 if args.generic else 1, graph=0 if args.no_graph else 1)
    plan = lib.plan(D, 1)
    nbytes = lib.query("cm21_find_bubbles_workspace_bytes", D, len(radii), 0)
    work = torch.empty(nbytes, dtype=torch.uint8, device="cuda")
    ion = torch.empty_like(upd)
    r_arr = np.ascontiguousarray(radii, np.float32)
    d_arr = np.ascontiguousarray(denoms, np.float32)
    torch.cuda.synchronize()
    stream = torch.cuda.Stream()  # graph capture needs a non-default stream

    def step():
        with torch.cuda.stream(stream):
            _step()

    def _step():
        lib.call("cm21_find_bubbles", plan.handle, lib.ptr(upd), lib.ptr(ion), r_arr.ctypes.data_as(lib.c_float_p),
                 d_arr.ctypes.data_as(lib.c_float_p), len(radii), prm, lib.ptr(work), nbytes,
                 lib.stream_handle(stream))

    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()
    # calibration: every stage instrumented, to find the dominant one
    lib.profile_collect()
    lib.profile_enable(True)
    for _ in range(2):
        step()
    torch.cuda.synchronize()
    lib.profile_enable(False)
    calib = lib.profile_collect()
    dom = max(calib, key=lambda k: calib[k][0])
    stages_all = {k: {"ms_per_launch": ms / n, "launches_per_box": n / 2} for k, (ms, n) in calib.items()}

    # timed region 1 (the reported value): K steps, no instrumentation; with
    # graph replay unless --no-graph
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        step()
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0
    ms_step = elapsed * 1e3 / args.steps
    boxes_s = args.steps / elapsed

    # timed region 2 (the roofline): the same K steps launched eagerly with HIP
    # events around every launch of the dominant stage, on the launch stream
    lib.profile_enable([dom])
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        step()
    torch.cuda.synchronize()
    ms_step_events = (time.perf_counter() - t0) * 1e3 / args.steps
    lib.profile_enable(False)
    prof = lib.profile_collect()
    x_HI = 1.0 - float(ion.double().mean())
    ab = alg_bytes(D)
    for k, v in stages_all.items():
        v["gbs_alg"] = (ab[k] / (v["ms_per_launch"] * 1e-3) / 1e9) if k in ab else None
    dom_ms = prof[dom][0] / prof[dom][1]
    achieved = ab[dom] / (dom_ms * 1e-3) / 1e9
    # calibrated rocprofv3 FETCH_SIZE + WRITE_SIZE per launch of the same
    # kernel and workload, from the committed counter summary
    traffic = None if args.generic else read_traffic(f"{dom}_D{D}_traffic.txt", "bytes_per_launch")
    # counter-measured VALU wave-instructions per launch of the same kernel
    valu = None if args.generic else read_traffic(f"{dom}_D{D}_traffic.txt", "valu_per_launch")

    # delta_T and P21(k) in 40 log bins k_fund..k_nyq on the result (once)
    def post():
        xH = 1.0 - ion
        dT, s = dt.brightness_temperature(xH, upd, 8.0)
        return dt.power_spectrum(dT, L, dt.log_edges(L, D, 40), field_sum=s)
    _, t_post = timed(post)
    _, t_post = timed(post)

    out = {
        "metric": "HII boxes/s at HII_DIM=256: plain vs batched vs surrogate @ |dx_HI|<0.01",
        "value": boxes_s,
        "unit": "boxes/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_step,
        "ms_per_step_eager_with_events": ms_step_events,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": "plain find_HII_bubbles, HII_DIM=256 nonlinear box from DIM=512 GRF + CIC "
                               "Zel'dovich move (BOX_LEN=300 Mpc, 40 radii, zeta=30, HII_FILTER=1)",
                   "HII_DIM": D, "DIM": DIM, "BOX_LEN": L, "n_radii": len(radii), "zeta": zeta,
                   "sigma_min2": smin, "growth": growth,
                   "path": ("plain-generic" if args.generic else "plain-fused") + ("" if args.no_graph else "+graph")},
        "roofline": {"bound": "hbm", "kernel": dom, "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": traffic,
                     "traffic_gbs": (traffic / (dom_ms * 1e-3) / 1e9) if traffic else None,
                     "traffic_frac": (traffic / (dom_ms * 1e-3) / 1e9 / HBM_PEAK_GBS) if traffic else None,
                     "alg_bytes_per_launch": ab[dom], "ms_per_launch": dom_ms,
                     "note": "achieved credits the unpruned algorithm's bytes; the pruned launch moves "
                             "`traffic` bytes (traffic_frac of peak) and is bound by VALU issue (valu_roofline)"},
        # what bounds the kernel: VALU issue (counter-measured instructions per
        # launch over this run's launch time)
        "valu_roofline": ({"achieved": valu / (dom_ms * 1e-3) / 1e9, "peak": VALU_PEAK_G,
                           "unit": "G wave64-instructions/s", "frac": valu / (dom_ms * 1e-3) / 1e9 / VALU_PEAK_G,
                           "valu_per_launch": valu,
                           "source": f"profiles/{dom}_D{D}_traffic.txt (from profiles/{dom}_D{D}_counters.md)"}
                          if valu else None),
        "model_roofline": {"B_alg_per_box": b_alg(D, len(radii), 1),
                           "frac": b_alg(D, len(radii), 1) * boxes_s / (HBM_PEAK_GBS * 1e9)},
        "stages": stages_all,
        "pipeline_ms": {"initialize_512_to_256": t_init, "perturb_cic": t_pert, "delta_T_and_P21": t_post},
        "x_HI": x_HI,
    }
    if not args.no_extra:
        del work, ion
        torch.cuda.empty_cache()
        out["extra"] = extra_workloads(torch, upd, args.steps)
        out["extra"]["config2_partly_ionized"]["vs_literal"] = out["extra"]["config2_partly_ionized"]["value"] / boxes_s
    if not args.no_batched:
        torch.cuda.empty_cache()
        out["batched"] = batched_bench(args, torch)
        out["batched"]["vs_plain"] = out["batched"]["value"] / boxes_s
        out["batched_partly_ionized"] = batched_bench(args, torch, partial=True)
        if "extra" in out:
            out["batched_partly_ionized"]["vs_plain_partly_ionized"] = (
                out["batched_partly_ionized"]["value"] / out["extra"]["config2_partly_ionized"]["value"])
    if not args.no_surrogate:
        torch.cuda.empty_cache()
        out["surrogate"] = surrogate_bench(args, torch, boxes_s)
    if not args.no_cpu_baseline:
        out["cpu_baseline"] = cpu_baseline(upd.cpu().numpy(), L, radii, denoms, zeta, deltac, fc_mean,
                                           args.cpu_budget,
                                           out.get("extra", {}).get("config1", {}).get("value"))
    print(json.dumps(out))


if __name__ == "__main__":
    main()
