"""Excursion-set ionization field (drop-in for Programs/find_HII_bubbles.py).

``conv_bubbles`` keeps the reference signature and return value
(find_HII_bubbles.py:17-49,192-193): a float32 (D,D,D) ``ionized`` box (0, 1,
or the partial zeta*f_coll painted on the last radius).  The host computes the
scalars the reference computes in Python (radius list, S(R), denominators,
barrier, mean collapsed fraction); the whole radius loop then runs on the GPU
through ``cm21_find_bubbles`` with no host synchronisation.

cosmolopy (``pb.sig_del``, ``pb.collapse_fraction``, find_HII_bubbles.py:57,60)
is not available.  The barrier and sigma_min are either given explicitly,
``deltac`` (or ``growth`` = D(z), giving deltac = 1.686/D) and ``sigma_min2``,
with S(R) from ``sig0`` (a callable) or by default from the synthetic
power-law spectrum of the box (cosmo_files.sigmas.PowerLawSpectrum); or, with
none of the three given, they follow from Z, INIT_PARAMS.ION_Tvir_MIN and the
Planck13 cosmology of INIT_PARAMS (cosmo_files.planck: sig_del, D(z), and
S(R) of the Planck13 spectrum), as the reference computes them.
"""
import functools
import importlib
import math
import os

import numpy as np
import torch
from scipy.special import erfc

from .. import _lib
from ..Parameter_files import INIT_PARAMS as P
from ..cosmo_files.planck import barrier as planck_barrier
from ..cosmo_files.sigmas import power_law_spectrum

DELTA_C_LINEAR = 1.686


def radius_schedule(box_len, hii_dim, py2_int_scale=False, r_max=None, r_min=None, delta_r_factor=None,
                    l_factor=None, cell=None):
    """float32 radii R_max, R_max/1.1, ... and the cell size of the stop rule.

    find_HII_bubbles.py:52-54 (scale: the caller's ``cell`` if given, else
    BoxSize/HIIdim, an integer division under Python 2 when py2_int_scale),
    :92, :123-125 (final_step when R/1.1 <= L_FACTOR*scale or <=
    R_BUBBLE_MIN, tested in float64), :125 (R cast to float32), :189 (R/1.1
    in float64).  The last radius returned is the final step.
    """
    r_max = P.R_BUBBLE_MAX if r_max is None else r_max
    r_min = P.R_BUBBLE_MIN if r_min is None else r_min
    fac = P.DELTA_R_FACTOR if delta_r_factor is None else delta_r_factor
    lf = float(np.float32(P.L_FACTOR if l_factor is None else l_factor))
    if cell is not None:
        scale = float(cell)
    else:
        scale = (int(box_len) // int(hii_dim)) if py2_int_scale else float(box_len) / float(hii_dim)
    R = float(r_max)
    radii = []
    while True:
        final = (R / fac) <= lf * scale or (R / fac) <= r_min
        R32 = np.float32(R)
        radii.append(R32)
        if final:
            break
        R = float(R32) / fac
    return np.array(radii, dtype=np.float32), scale


def barrier_denominators(radii, sigma_min2, sig0_fn):
    """denom = float32(sqrt(2 (smin - float32(S(R))))) per radius
    (find_HII_bubbles.py:126-128); raises ValueError where smin <= S(R)."""
    out = np.empty(len(radii), dtype=np.float32)
    for i, R in enumerate(radii):
        S0 = np.float32(sig0_fn(float(R)))
        arg = 2.0 * (float(sigma_min2) - float(S0))
        if not arg > 0.0:
            raise ValueError(f"sigma_min^2={sigma_min2} <= S(R={float(R)})={float(S0)}: no barrier at this radius")
        out[i] = np.float32(math.sqrt(arg))
    return out


@functools.lru_cache(maxsize=16)
def _default_spectrum(D, box_len):
    return power_law_spectrum(D, box_len)


def _default_sig0(D, box_len):
    # one spectrum per box shape, so S(R) of its radii is computed once
    return _default_spectrum(int(D), float(box_len)).sig0_discrete


def bubble_setup(D, box_len, scale, deltac, sigma_min2, sig0_fn, py2_int_scale=False):
    """Host-side scalars of conv_bubbles: radii, denominators, fc_mean."""
    radii, cell = radius_schedule(box_len, D, py2_int_scale, cell=scale)
    lf = float(np.float32(P.L_FACTOR))
    if P.R_BUBBLE_MIN < lf * cell:  # find_HII_bubbles.py:115-120
        if 2.0 * (float(sigma_min2) - sig0_fn(lf * cell)) < 0:
            raise ValueError(f"sigma_min^2={sigma_min2} < S(L_FACTOR*cell={lf * cell})")
    denoms = barrier_denominators(radii, sigma_min2, sig0_fn)
    fc_mean = np.float32(erfc(float(deltac) / (math.sqrt(2.0) * math.sqrt(float(sigma_min2)))))
    return radii, denoms, fc_mean


def conv_bubbles(deltax, param_dict, Z=None, scale=None, fil=1, update=0, LE=False, visualize=0, quiet=False,
                 *, deltac=None, growth=None, sigma_min2=None, zeta=None, box_len=None, sig0=None,
                 strict_filter=False, py2_int_scale=False, sphere_periodic=False, sphere_in_cells=False,
                 fused=True, graph=False, stream=None):
    """Ionized box from a density box by the excursion-set radius loop.

    deltax: float32 (D,D,D) NumPy array or CUDA tensor (returned in the same
    kind).  param_dict: needs 'BoxSize' and 'HIIdim' (boxio.parse_filename).
    fil: 0 real-space top-hat, 1 k-space top-hat, 2 Gaussian (reference
    fall-through chain unless strict_filter).  update: 0 centre pixel,
    1 sphere paint (R in Mpc used as a cell radius, clipped at the faces, as the
    reference; sphere_in_cells/sphere_periodic give R/cell and wrap-around).
    LE: True for linear z=0 input (no growth scaling, no normalisation).
    fused: use the fused, pruned transform kernels where supported (centre-pixel
    update, filter 1 or 2, D in 64/96/128/160/256/512); False forces rocFFT plus
    separate filter and threshold kernels.
    graph: capture the loop into a HIP graph and replay it when called again
    with the same tensors and parameters (needs a non-default stream).
    """
    if visualize:
        raise ValueError("visualize is not supported by the GPU path")
    _lib.require_gpu()
    D = int(param_dict.get("HIIdim", np.shape(deltax)[0]))
    L = float(box_len if box_len is not None else param_dict["BoxSize"])
    if np.shape(deltax) != (D, D, D):
        raise ValueError(f"deltax shape {np.shape(deltax)} does not match HIIdim={D}")
    if deltac is None and growth is None and sigma_min2 is None:
        # Planck-cosmology mode (find_HII_bubbles.py:57-60 without cosmolopy):
        # the barrier from ION_Tvir_MIN at Z and S(R) of the Planck13 spectrum
        z = Z if Z is not None else param_dict.get("Z")
        if z is None:
            raise ValueError("give Z (or param_dict['Z']) for the Planck mode, or deltac/growth and sigma_min2")
        cos = planck_barrier(float(z))
        deltac, sigma_min2 = cos["deltac"], cos["sigma_min2"]
        if sig0 is None:
            sig0 = cos["sig0"]
    if deltac is None:
        if growth is None:
            raise ValueError("give deltac or growth with sigma_min2, or neither (Planck mode)")
        deltac = DELTA_C_LINEAR / float(growth)
    if sigma_min2 is None:
        raise ValueError("give sigma_min2 with deltac or growth, or neither (Planck mode)")
    zeta = float(P.ZETA if zeta is None else zeta)
    fgrowth = np.float32(float(deltac) / DELTA_C_LINEAR)
    sig0_fn = sig0 if sig0 is not None else _default_sig0(D, L)
    radii, denoms, fc_mean = bubble_setup(D, L, scale, deltac, sigma_min2, sig0_fn, py2_int_scale)
    if not quiet:
        print(f"Using filter_type {fil}")
        for R, d in zip(radii, denoms):
            print(f"R={R} Mpc denom {d}")
    cell = float(L) / D
    prm = _lib.BubbleParams(
        box_len=L, deltac=float(np.float32(deltac)), zeta=zeta, growth_inv=float(fgrowth),
        fc_mean=float(fc_mean), linear_input=1 if LE else 0, filter_type=int(fil),
        ref_fallthrough=0 if strict_filter else 1, update=int(update),
        sphere_periodic=1 if sphere_periodic else 0,
        sphere_r_scale=(1.0 / cell) if sphere_in_cells else 1.0, fused=1 if fused else 0,
        graph=1 if graph else 0)
    is_tensor = isinstance(deltax, torch.Tensor)
    dev = torch.device("cuda")
    caller = torch.cuda.current_stream()
    s = caller if stream is None else stream
    plan = _lib.plan(D, 1)
    nbytes = _lib.query("cm21_find_bubbles_workspace_bytes", D, len(radii), int(update))
    r = np.ascontiguousarray(radii, dtype=np.float32)
    dn = np.ascontiguousarray(denoms, dtype=np.float32)
    # Everything runs on `s`, ordered after the caller's pending work and
    # before the caller's later work: the input upload, the loop, the read-back.
    if s != caller:
        s.wait_stream(caller)
    with torch.cuda.stream(s):
        d = deltax if is_tensor else torch.from_numpy(np.ascontiguousarray(deltax, dtype=np.float32))
        d = d.to(dev, dtype=torch.float32).contiguous()
        if graph:
            # a replayed graph reuses the captured addresses: keep the buffers
            ion, work = _graph_buffers(d, nbytes, s)
        else:
            ion = torch.empty_like(d)
            work = torch.empty(nbytes, dtype=torch.uint8, device=dev)
        _lib.call("cm21_find_bubbles", plan.handle, _lib.ptr(d), _lib.ptr(ion),
                  r.ctypes.data_as(_lib.c_float_p), dn.ctypes.data_as(_lib.c_float_p), len(radii),
                  prm, _lib.ptr(work), nbytes, _lib.stream_handle(s))
        out = ion.clone() if (graph and is_tensor) else ion
        host = None if is_tensor else out.cpu().numpy()
    if s != caller:
        caller.wait_stream(s)
        for t in (d, out, work):
            t.record_stream(caller)
    return out if is_tensor else host


# Graph replay needs the captured buffers to stay alive.  One entry is kept per
# (box side, stream), the granularity at which the library keeps graphs: a new
# input address replaces the entry (and its buffers) of its shape and stream.
_graph_cache = {}


def _graph_buffers(d, nbytes, stream):
    """Output and workspace kept alive for graph replay of the input `d`."""
    slot = (tuple(d.shape), nbytes, stream.cuda_stream)
    cur = _graph_cache.get(slot)
    if cur is None or cur[0] != d.data_ptr():
        _graph_cache[slot] = (d.data_ptr(), torch.empty_like(d),
                              torch.empty(nbytes, dtype=torch.uint8, device=d.device))
    return _graph_cache[slot][1], _graph_cache[slot][2]


def xH_filename(Z, nf, fil, dim, box_size):
    """Output name of find_HII_bubbles.py:227-228."""
    return ("xH_nohalos_z{0:.2f}_nf{1:f}_eff{2:.1f}_HIIfilter{3:d}_Tvirmin{4:.1f}_RHIImax{5:d}_{6:d}_{7:d}Mpc.npy"
            .format(Z, nf, P.ZETA, int(fil), P.ION_Tvir_MIN, int(P.R_BUBBLE_MAX), dim, int(box_size)))


def run(opts, args):
    """CLI driver of find_HII_bubbles.py:195-232: load, pick LE from the file
    name ('updated*' = nonlinear input at its z, else linear input extrapolated
    to opts.Z), run, save 1 - ionized under the reference's name."""
    boxio = importlib.import_module(__package__.rsplit(".", 1)[0] + ".tocmfastpy.boxio")
    if len(args) == 0:
        raise ValueError("Must specify density box input!")
    fname = args[0]
    if fname.endswith("npy"):
        param_dict = boxio.parse_filename(fname)
        d1 = np.load(fname)
    else:
        b1 = boxio.readbox(fname)
        d1 = b1.box_data[:256, :256, :256]
        param_dict = b1.param_dict
    if not param_dict["filename"].startswith("updated"):
        Z = float(opts.Z)
        LE = True
    else:
        Z = param_dict["z"]
        LE = False
    growth = getattr(opts, "GROWTH", None)
    smin2 = getattr(opts, "SIGMA_MIN2", None)
    if (growth is None) != (smin2 is None):
        raise ValueError("give both opts.GROWTH and opts.SIGMA_MIN2, or neither (Planck mode)")
    extra = {} if growth is None else {"growth": float(growth), "sigma_min2": float(smin2)}
    ion = conv_bubbles(d1, param_dict, Z, fil=int(opts.FILTER_TYPE), update=int(opts.UPDATE_TYPE), LE=LE,
                       quiet=bool(getattr(opts, "QUIET", False)), **extra)
    nf = 1 - float(np.mean(ion, dtype=np.float64))
    outname = xH_filename(Z, nf, opts.FILTER_TYPE, ion.shape[0], param_dict["BoxSize"])
    np.save(os.path.join(opts.DIR, outname), 1 - ion)
    return ion


def main(argv=None):
    import argparse
    o = argparse.ArgumentParser(description="excursion-set ionization field on the GPU")
    o.add_argument("box")
    o.add_argument("-d", "--dir", dest="DIR", default=os.path.join(P.IO_DIR, "Boxes"))
    o.add_argument("-f", "--filt", dest="FILTER_TYPE", default=P.HII_FILTER, type=int)
    o.add_argument("-u", "--upd", dest="UPDATE_TYPE", default=P.FIND_BUBBLE_ALGORITHM, type=int)
    o.add_argument("-z", "--reds", dest="Z", default=12.0, type=float)
    o.add_argument("-g", "--growth", dest="GROWTH", type=float, default=None,
                   help="linear growth D(z); with -s, instead of the Planck13 barrier")
    o.add_argument("-s", "--sigma-min2", dest="SIGMA_MIN2", type=float, default=None)
    o.add_argument("-q", "--qui", dest="QUIET", action="store_true")
    opts = o.parse_args(argv)
    return run(opts, [opts.box])


if __name__ == "__main__":
    main()
