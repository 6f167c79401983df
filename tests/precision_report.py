"""Precision of find_HII_bubbles at HII_DIM=512 against the FP64 oracle.

Workload: a 512^3 linear GRF box (BOX_LEN = 500 Mpc, seed 0, P(k) ~ k^-2.5
exp(-(k/k_nyq)^2), per-cell sigma 0.6), 41 radii, zeta = 25, HII_FILTER 1,
barrier 1.686.  Two values of sigma_min^2: 4.0 (the literal parameters,
every cell ionizes at the first radius) and 0.5 (partly ionized).  The GPU
result (the fused path through the C ABI) is compared with the FP64 oracle
(oracle/bubbles.find_bubbles_f64, transforms on the host cores) by the
flipped-cell fraction (all cells and outside the 1e-5 threshold band), the
relative error of x_HI, and the per-bin relative error of P21(k) (40 log bins
k_fund..k_nyq, the same estimator on both fields).

Not collected by pytest (it runs for minutes); run on a GPU host as
    python tests/precision_report.py [--out report.md]
"""
import argparse
import importlib
import os
import sys
import time

import numpy as np
import torch

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)

from oracle import bubbles, sigma  # noqa: E402


def pkg(name):
    return importlib.import_module("21cmpy_amd." + name)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--D", type=int, default=512)
    ap.add_argument("--box-len", type=float, default=500.0)
    ap.add_argument("--zeta", type=float, default=25.0)
    ap.add_argument("--smin", type=float, nargs="+", default=[4.0, 0.5])
    ap.add_argument("--workers", type=int, default=min(16, len(os.sched_getaffinity(0))))
    ap.add_argument("--out", default=None)
    args = ap.parse_args()
    g = pkg("gpu")
    fb = pkg("Programs.find_HII_bubbles")
    dt = pkg("Programs.delta_T")
    sig = pkg("cosmo_files.sigmas")
    D, L = args.D, args.box_len
    sp = sig.PowerLawSpectrum(D, L)
    delta_d = g.c2r(g.grf_init(D, L, sp.amp, 0))
    delta = delta_d.cpu().numpy()
    radii, _ = sigma.radius_schedule(L, D)
    edges = dt.log_edges(L, D, 40)
    lines = [f"# find_HII_bubbles precision at {D}^3 against the FP64 oracle", "",
             f"Command: `python tests/precision_report.py` ({args.workers} host threads for the oracle's "
             f"transforms).  BOX_LEN {L:g} Mpc, {len(radii)} radii, zeta {args.zeta:g}, HII_FILTER 1, "
             "barrier 1.686, linear input, GRF seed 0.", "",
             "| sigma_min^2 | x_HI GPU | x_HI FP64 | rel. err x_HI | flipped cells | flipped outside band "
             "| cells in band | max P21 rel. err | median P21 rel. err | GPU s | oracle s |",
             "|---|---|---|---|---|---|---|---|---|---|---|"]
    for smin in args.smin:
        print(f"sigma_min^2 = {smin}: GPU", flush=True)
        t0 = time.perf_counter()
        got = fb.conv_bubbles(delta_d, {"BoxSize": L, "HIIdim": D}, LE=True, quiet=True, deltac=1.686,
                              sigma_min2=smin, zeta=args.zeta, sig0=sp.sig0_discrete)
        torch.cuda.synchronize()
        t_gpu = time.perf_counter() - t0
        print(f"  oracle ({args.workers} threads)", flush=True)
        den = sigma.denominators(radii, smin, sp.sig0_discrete)
        t0 = time.perf_counter()
        want, margin = bubbles.find_bubbles_f64(delta, L, radii, den, args.zeta, 1.686, LE=True,
                                                fft_workers=args.workers)
        t_cpu = time.perf_counter() - t0
        g_np = got.cpu().numpy()
        fg, fw = g_np == 1, want == 1
        flips = float(np.mean(fg != fw))
        band = margin <= 1e-5
        flips_out = float(np.mean((fg != fw) & ~band))
        xg = 1.0 - float(np.mean(g_np, dtype=np.float64))
        xw = 1.0 - float(np.mean(want))
        rel = abs(xg - xw) / xw if xw > 0 else (0.0 if xg == xw else float("inf"))
        pks = []
        for f in (got, torch.from_numpy(want.astype(np.float32)).cuda()):
            dT, s = dt.brightness_temperature((1.0 - f).contiguous(), delta_d, 8.0)
            _, _, pk, _ = dt.power_spectrum(dT, L, edges, field_sum=s)
            pks.append(pk)
        ok = pks[1] > 1e-3 * max(pks[1].max(), 1e-300)
        if ok.any():
            e = np.abs(pks[0][ok] - pks[1][ok]) / pks[1][ok]
            pmax, pmed = f"{e.max():.2e}", f"{np.median(e):.2e}"
        else:
            pmax = pmed = "n/a (no 21-cm signal: fully ionized)"
        lines.append(f"| {smin:g} | {xg:.6f} | {xw:.6f} | {rel:.2e} | {flips:.2e} | {flips_out:.2e} | "
                     f"{np.mean(band):.2e} | {pmax} | {pmed} | {t_gpu:.2f} | {t_cpu:.1f} |")
        print(lines[-1], flush=True)
        del got, want, margin
    text = "\n".join(lines) + "\n"
    print(text)
    if args.out:
        with open(args.out, "w") as f:
            f.write(text)


if __name__ == "__main__":
    main()
