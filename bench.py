"""Benchmark of the find_HII_bubbles hot path on one MI355X.

Workload (the full plain pipeline at HII_DIM=256): a DIM=512, BOX_LEN=300 Mpc
Gaussian random field (P(k) ~ k^-2.5 exp(-(k/k_nyq)^2), per-cell sigma 0.6,
seed 0) is smoothed and nearest-picked to 256^3, moved by the Zel'dovich
displacement at growth 0.12 (D_init = 0.12*9/301) with a cloud-in-cell
deposit, and the resulting nonlinear 256^3 box goes through find_HII_bubbles:
40 radii, zeta = 30, k-space top-hat (HII_FILTER = 1), box-mean
normalisation, sigma_min^2 = 4.0, barrier 1.686/0.12.

A step is one find_HII_bubbles pass over that box, with the box already in
HBM.  The JSON line reports boxes/s, the roofline of the dominant stage
(algorithmic bytes per launch over its HIP-event-timed average duration) and a
CPU baseline: the NumPy FP32 oracle of the same algorithm on a bounded sample.
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


def alg_bytes(D, norm=1):
    """Algorithmic bytes per launch (one radius) of each stage, from the unit
    model B_alg = 4N + 8Nc + n_R (24 Nc + (16 + 4 norm) N): filter 16Nc,
    c2r 8Nc + 4N, f_coll box sum 4N, threshold 12N.  A fused kernel carries the
    model bytes of the stages it replaces: columns = filter; lines = c2r plus
    the box sum (nonlinear input) or plus the threshold (linear input); rows
    are the middle of the c2r and carry none of their own."""
    N = D ** 3
    Nc = D * D * (D // 2 + 1)
    return {
        "r2c": 4 * N + 8 * Nc,
        "filter": 16 * Nc,
        "c2r": 8 * Nc + 4 * N,
        "fcoll_sum": 4 * N,
        "threshold": 12 * N,
        "fused_columns": 16 * Nc,
        "fused_rows": 0,
        "fused_lines": 8 * Nc + 4 * N + (4 * N if norm else 12 * N),
    }


def b_alg(D, n_r, norm):
    N = D ** 3
    Nc = D * D * (D // 2 + 1)
    return 4 * N + 8 * Nc + n_r * (24 * Nc + (16 + 4 * norm) * N)


def cpu_baseline(delta, L, radii, denoms, zeta, deltac, fc_mean, budget_s=15.0):
    """NumPy FP32 restatement (the oracle) on the same box: time the forward
    transform plus as many radii as fit the budget; extrapolate to one box."""
    from oracle import bubbles
    D = delta.shape[0]
    t0 = time.perf_counter()
    d = (delta * np.float32(deltac / 1.686)).astype(np.float32)
    spec = np.fft.rfftn(d.astype(np.float64)).astype(np.complex64)
    km = bubbles.kmag_half_f32(D, bubbles.delta_k_f32(L))
    t_fwd = time.perf_counter() - t0
    done, t_rad = 0, 0.0
    ion = np.zeros(delta.shape, np.float32)
    for R, den in zip(radii, denoms):
        t1 = time.perf_counter()
        w0, cut, w2 = bubbles.window_f32(km, R, 1, True)
        filt = bubbles.apply_window_f32(spec, w0, cut, w2)
        sm = np.fft.irfftn(filt.astype(np.complex128), s=(D, D, D), axes=(0, 1, 2)).astype(np.float32)
        fc = bubbles.fcoll_f32(sm, deltac, den)
        fc = (fc * np.float32(float(fc_mean) / (fc.astype(np.float64).sum() / D ** 3))).astype(np.float32)
        ion[fc.astype(np.float64) >= 1.0 / zeta] = 1.0
        t_rad += time.perf_counter() - t1
        done += 1
        if t_fwd + t_rad > budget_s:
            break
    per_box = t_fwd + t_rad / done * len(radii)
    return {"value": 1.0 / per_box, "unit": "boxes/s", "cores": 1, "kind": "port",
            "sample": f"NumPy FP32 oracle (oracle/bubbles.py, pocketfft single thread) on the same {D}^3 box: "
                      f"forward transform + {done} of {len(radii)} radii timed ({t_fwd + t_rad:.1f} s), "
                      f"extrapolated to one box"}


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


def batched_bench(args, torch):
    """The 64-box batch (mixed HII_DIM 96/128/160, per-box zeta, z, sigma_min^2
    from the committed manifest; linear z=0 GRF boxes, cell 1.5625 Mpc, 37
    radii, HII_FILTER 1) through cm21_find_bubbles_batched as one graph.
    Returns the result object for the JSON line."""
    bm = importlib.import_module("21cmpy_amd.Programs.find_HII_bubbles_batched")
    g = importlib.import_module("21cmpy_amd.gpu")
    sig = importlib.import_module("21cmpy_amd.cosmo_files.sigmas")
    lib = importlib.import_module("21cmpy_amd._lib")
    specs = bm.load_manifest()["boxes"]
    stream = torch.cuda.Stream()
    with torch.cuda.stream(stream):
        bb = bm.BatchedBubbles(specs, fil=1, LE=True, graph=not args.no_graph, stream=stream)
        amps = {}
        for i, s in enumerate(specs):
            key = (s["D"], s["box_len"])
            if key not in amps:
                amps[key] = sig.PowerLawSpectrum(*key).amp
            bb.set_delta(i, g.c2r(g.grf_init(s["D"], s["box_len"], amps[key], s["seed"])))
        tuned = None if args.no_tune else bb.autotune(reps=3, save_path=args.save_tiles)
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
    return {
        "value": value, "unit": "boxes/s", "boxes_per_batch": nb, "steps": steps,
        "ms_per_batch": elapsed * 1e3 / steps,
        "workload": "64 linear GRF boxes, HII_DIM in {96,128,160} (21/25/18), BOX_LEN = 1.5625 Mpc x HII_DIM, "
                    "37 radii, per-box zeta in [10,60], z in {7..10} (deltac = 1.686/D(z), D(z)=0.12*9/(1+z)), "
                    "sigma_min^2 in [2,6], HII_FILTER 1; one graph per batch",
        "model_roofline": {"B_alg_per_batch": balg, "frac": balg * steps / elapsed / (HBM_PEAK_GBS * 1e9)},
        "stages_one_batch_eager": stages,
        "tiles": ({k: v["chosen"] for k, v in tuned.items()} if tuned else
                  {k: {"columns": c, "lines": l} for k, (c, l) in bb.tiles().items()}),
        "x_HI_mean": float(np.mean(xhi)), "x_HI_min": float(np.min(xhi)),
    }


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


def plain_rate(torch, delta, L, deltac, smin, zeta, steps, linear=True):
    """Boxes/s of the plain single-box loop (fused path, graph replay) on one
    resident box, radii and denominators computed beforehand."""
    lib = importlib.import_module("21cmpy_amd._lib")
    fb = importlib.import_module("21cmpy_amd.Programs.find_HII_bubbles")
    sig = importlib.import_module("21cmpy_amd.cosmo_files.sigmas")
    D = delta.shape[0]
    radii, denoms, fc_mean = fb.bubble_setup(D, L, None, deltac, smin, sig.PowerLawSpectrum(D, L).sig0_discrete)
    prm = lib.BubbleParams(box_len=L, deltac=float(np.float32(deltac)), zeta=zeta,
                           growth_inv=float(np.float32(deltac / 1.686)), fc_mean=float(fc_mean),
                           linear_input=1 if linear else 0, filter_type=1, ref_fallthrough=1, update=0,
                           sphere_periodic=0, sphere_r_scale=1.0, fused=1, graph=1)
    plan = lib.plan(D, 1)
    nbytes = lib.query("cm21_find_bubbles_workspace_bytes", D, len(radii), 0)
    work = torch.empty(nbytes, dtype=torch.uint8, device="cuda")
    ion = torch.empty_like(delta)
    r_arr = np.ascontiguousarray(radii, np.float32)
    d_arr = np.ascontiguousarray(denoms, np.float32)
    stream = torch.cuda.Stream()

    def step():
        lib.call("cm21_find_bubbles", plan.handle, lib.ptr(delta), lib.ptr(ion), r_arr.ctypes.data_as(lib.c_float_p),
                 d_arr.ctypes.data_as(lib.c_float_p), len(radii), prm, lib.ptr(work), nbytes,
                 lib.stream_handle(stream))

    torch.cuda.synchronize()
    for _ in range(3):
        step()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        step()
    torch.cuda.synchronize()
    return steps / (time.perf_counter() - t0)


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
    # plain method on the same box size, one box per call through the C ABI
    # (graph replay), host set-up outside the timed region
    s0 = specs[0]
    plain128 = plain_rate(torch, boxes.delta[0].contiguous(), L, s0["deltac"], s0["sigma_min2"], s0["zeta"],
                          steps)
    mf, other = unet_flops(D)
    achieved = B * (mf + other) / (fwd_ms * 1e-3) / 1e12
    out = {
        "value": value, "unit": "boxes/s", "HII_DIM": D, "BOX_LEN": L, "batch": B, "steps": steps,
        "ms_per_forward_batch": fwd_ms, "ms_per_train_step_batch": train_ms,
        "trained": trained, "weights_meta": meta,
        "accuracy_vs_plain_heldout16": summary,
        "plain_same_HII_DIM_boxes_s": plain128,
        "roofline": {"bound": "mfma", "achieved": achieved, "peak": 2500.0, "unit": "TFLOP/s",
                     "frac": achieved / 2500.0, "flops_per_box": mf + other, "mfma_flops_per_box": mf},
        "workload": "3-level 16/32/64 3D U-Net, bf16 MFMA convs + fp32 accumulation, on linear GRF boxes "
                    "(zeta in [10,60], sigma_min^2 in [2,6], barrier 4.0)",
    }
    if plain_boxes_s:
        out["vs_plain_256"] = value / plain_boxes_s
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
    ap.add_argument("--mode", choices=["bench", "loop", "copy", "batched", "surrogate"], default="bench",
                    help="loop: only the radius loop on a GRF box (for counter runs); copy: a 512 MiB device "
                         "copy (to calibrate FETCH_SIZE/WRITE_SIZE)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: this build measures one GPU; --gpus {args.gpus} is not supported "
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
    traffic = None  # calibrated rocprofv3 FETCH_SIZE + WRITE_SIZE per launch, when a summary exists
    tfile = os.path.join(ROOT, "profiles", f"{dom}_D{D}_traffic.txt")
    if os.path.exists(tfile) and not args.generic:
        for line in open(tfile):
            if line.startswith("bytes_per_launch:"):
                traffic = float(line.split(":", 1)[1])

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
                     "alg_bytes_per_launch": ab[dom], "ms_per_launch": dom_ms},
        "model_roofline": {"B_alg_per_box": b_alg(D, len(radii), 1),
                           "frac": b_alg(D, len(radii), 1) * boxes_s / (HBM_PEAK_GBS * 1e9)},
        "stages": stages_all,
        "pipeline_ms": {"initialize_512_to_256": t_init, "perturb_cic": t_pert, "delta_T_and_P21": t_post},
        "x_HI": x_HI,
    }
    if not args.no_batched:
        del work, ion
        torch.cuda.empty_cache()
        out["batched"] = batched_bench(args, torch)
        out["batched"]["vs_plain"] = out["batched"]["value"] / boxes_s
    if not args.no_surrogate:
        torch.cuda.empty_cache()
        out["surrogate"] = surrogate_bench(args, torch, boxes_s)
    if not args.no_cpu_baseline:
        out["cpu_baseline"] = cpu_baseline(upd.cpu().numpy(), L, radii, denoms, zeta, deltac, fc_mean,
                                           args.cpu_budget)
        out["cpu_baseline"]["cores_available"] = len(os.sched_getaffinity(0))
    print(json.dumps(out))


if __name__ == "__main__":
    main()
