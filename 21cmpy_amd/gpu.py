"""Typed wrappers of the cm21 C ABI on torch CUDA tensors.

Each function validates shapes and dtypes on the host before launching (the
kernels trust their arguments), allocates outputs and workspaces with torch,
and runs asynchronously on the current torch stream.  Spectra are complex64
tensors of shape (D, D, D//2+1); real boxes are float32 (D, D, D).
"""
import ctypes

import numpy as np
import torch

from . import _lib

_DEV = torch.device("cuda")


def _check_box(t, D, name, dtype=torch.float32):
    if not isinstance(t, torch.Tensor) or not t.is_cuda:
        raise TypeError(f"{name} must be a CUDA tensor")
    if t.dtype != dtype or tuple(t.shape) != (D, D, D) or not t.is_contiguous():
        raise ValueError(f"{name}: expected contiguous {dtype} {(D, D, D)}, got {t.dtype} {tuple(t.shape)}")


def _check_spec(t, D, name):
    if not isinstance(t, torch.Tensor) or not t.is_cuda:
        raise TypeError(f"{name} must be a CUDA tensor")
    want = (D, D, D // 2 + 1)
    if t.dtype != torch.complex64 or tuple(t.shape) != want or not t.is_contiguous():
        raise ValueError(f"{name}: expected contiguous complex64 {want}, got {t.dtype} {tuple(t.shape)}")


def empty_spec(D):
    return torch.empty((D, D, D // 2 + 1), dtype=torch.complex64, device=_DEV)


def delta_k(box_len):
    """DELTA_K = float32(2 pi / float32(BOX_LEN)) (INIT_PARAMS.py:10)."""
    return float(np.float32(2.0 * np.pi / float(np.float32(box_len))))


def r2c(x, out=None):
    D = x.shape[0]
    _check_box(x, D, "x")
    out = empty_spec(D) if out is None else out
    _check_spec(out, D, "out")
    _lib.call("cm21_r2c", _lib.plan(D).handle, _lib.ptr(x), _lib.ptr(out), _lib.stream_handle())
    return out


def c2r(spec, out=None, preserve_input=True):
    """Unnormalised inverse real transform (the input is copied first unless
    preserve_input is False, because the transform may overwrite it)."""
    D = spec.shape[0]
    _check_spec(spec, D, "spec")
    src = spec.clone() if preserve_input else spec
    out = torch.empty((D, D, D), dtype=torch.float32, device=_DEV) if out is None else out
    _check_box(out, D, "out")
    _lib.call("cm21_c2r", _lib.plan(D).handle, _lib.ptr(src), _lib.ptr(out), _lib.stream_handle())
    return out


def hii_filter(spec, box_len, R, filter_type=1, ref_fallthrough=True, korder=0, scale=1.0, out=None):
    D = spec.shape[0]
    _check_spec(spec, D, "spec")
    out = torch.empty_like(spec) if out is None else out
    _check_spec(out, D, "out")
    _lib.call("cm21_hii_filter", _lib.ptr(out), _lib.ptr(spec), D, delta_k(box_len), int(filter_type),
              1 if ref_fallthrough else 0, int(korder), float(R), float(scale), _lib.stream_handle())
    return out


def fcoll_sum(smoothed, deltac, denom):
    D = smoothed.shape[0]
    _check_box(smoothed, D, "smoothed")
    out = torch.empty(1, dtype=torch.float64, device=_DEV)
    work = torch.empty(_lib.query("cm21_reduce_workspace_bytes"), dtype=torch.uint8, device=_DEV)
    _lib.call("cm21_fcoll_sum", _lib.ptr(smoothed), D, float(deltac), float(denom), _lib.ptr(out), _lib.ptr(work),
              _lib.stream_handle())
    return out


def fcoll_threshold(ionized, smoothed, deltac, denom, zeta, dev_sum=None, fc_mean=0.0, final_step=False):
    D = smoothed.shape[0]
    _check_box(smoothed, D, "smoothed")
    _check_box(ionized, D, "ionized")
    _lib.call("cm21_fcoll_threshold", _lib.ptr(ionized), _lib.ptr(smoothed), D, float(deltac), float(denom),
              float(zeta), _lib.ptr(dev_sum) if dev_sum is not None else None, float(fc_mean),
              1 if final_step else 0, _lib.stream_handle())
    return ionized


def update_sphere(ionized, smoothed, deltac, denom, zeta, R_cells, periodic=False, dev_sum=None, fc_mean=0.0):
    D = smoothed.shape[0]
    _check_box(smoothed, D, "smoothed")
    _check_box(ionized, D, "ionized")
    work = torch.empty(_lib.query("cm21_sphere_workspace_bytes", D), dtype=torch.uint8, device=_DEV)
    _lib.call("cm21_update_sphere", _lib.ptr(ionized), _lib.ptr(smoothed), D, float(deltac), float(denom),
              float(zeta), _lib.ptr(dev_sum) if dev_sum is not None else None, float(fc_mean), float(R_cells),
              1 if periodic else 0, _lib.ptr(work), _lib.stream_handle())
    return ionized


def grf_init(D, box_len, amp, seed, box_id=0, n_index=-2.5, k_cut=None, out=None):
    out = empty_spec(D) if out is None else out
    _check_spec(out, D, "out")
    k_cut = np.pi * D / float(box_len) if k_cut is None else float(k_cut)
    _lib.call("cm21_grf_init", _lib.ptr(out), int(D), float(box_len), float(amp), float(n_index), k_cut,
              ctypes.c_uint32(int(seed) & 0xFFFFFFFF), ctypes.c_uint32(int(box_id) & 0xFFFFFFFF),
              _lib.stream_handle())
    return out


def grf_init_table(D, box_len, k_table, p_table, seed, box_id=0, out=None):
    """GRF half spectrum with P(k) interpolated from a (K, P) table
    (float32, K increasing), as init_kernel does (initialize.cu:58-95)."""
    out = empty_spec(D) if out is None else out
    _check_spec(out, D, "out")
    K = torch.as_tensor(np.ascontiguousarray(k_table, dtype=np.float32)).to(_DEV)
    Pt = torch.as_tensor(np.ascontiguousarray(p_table, dtype=np.float32)).to(_DEV)
    if K.ndim != 1 or K.shape != Pt.shape or not 2 <= K.numel() <= 2048:
        raise ValueError("k_table and p_table must be 1-D float32 arrays of equal length in [2, 2048]")
    if not bool(torch.all(K[1:] > K[:-1])):
        raise ValueError("k_table must be strictly increasing")
    _lib.call("cm21_grf_init_table", _lib.ptr(out), int(D), float(box_len), _lib.ptr(K), _lib.ptr(Pt), int(K.numel()),
              ctypes.c_uint32(int(seed) & 0xFFFFFFFF), ctypes.c_uint32(int(box_id) & 0xFFFFFFFF),
              _lib.stream_handle())
    return out


def subsample(large, sw, out=None):
    D = large.shape[0]
    _check_box(large, D, "large")
    out = torch.empty((sw, sw, sw), dtype=torch.float32, device=_DEV) if out is None else out
    _check_box(out, sw, "out")
    pf = float(np.float32(D) / np.float32(sw))
    _lib.call("cm21_subsample", _lib.ptr(out), _lib.ptr(large), D, int(sw), pf, _lib.stream_handle())
    return out


def velocity(spec, box_len, comp, filter_type=0, R=0.0, out=None):
    D = spec.shape[0]
    _check_spec(spec, D, "spec")
    out = torch.empty_like(spec) if out is None else out
    _check_spec(out, D, "out")
    _lib.call("cm21_velocity", _lib.ptr(out), _lib.ptr(spec), D, delta_k(box_len), int(comp), int(filter_type),
              float(R), _lib.stream_handle())
    return out


def move_mass(deltax, vx, vy, vz, disp_scale, growth_init, scheme="cic", out=None):
    DIM = deltax.shape[0]
    sw = vx.shape[0]
    _check_box(deltax, DIM, "deltax")
    for name, v in (("vx", vx), ("vy", vy), ("vz", vz)):
        _check_box(v, sw, name)
    out = torch.empty((sw, sw, sw), dtype=torch.float32, device=_DEV) if out is None else out
    _check_box(out, sw, "out")
    sch = {"ngp": 0, "cic": 1}[scheme]
    _lib.call("cm21_move_mass", _lib.ptr(out), _lib.ptr(deltax), _lib.ptr(vx), _lib.ptr(vy), _lib.ptr(vz), DIM, sw,
              float(disp_scale), float(growth_init), sch, _lib.stream_handle())
    return out


def delta_T(xH, delta, tb_const, out=None):
    D = xH.shape[0]
    _check_box(xH, D, "xH")
    _check_box(delta, D, "delta")
    out = torch.empty_like(xH) if out is None else out
    s = torch.empty(1, dtype=torch.float64, device=_DEV)
    work = torch.empty(_lib.query("cm21_reduce_workspace_bytes"), dtype=torch.uint8, device=_DEV)
    _lib.call("cm21_delta_T", _lib.ptr(out), _lib.ptr(xH), _lib.ptr(delta), D, float(tb_const), _lib.ptr(s),
              _lib.ptr(work), _lib.stream_handle())
    return out, s


def bin_lut(D, box_len, edges):
    """int32 table: bin of each integer |n|^2 (first edge >= dk sqrt(|n|^2);
    -1 for the DC mode, nbins above the last edge)."""
    m = np.arange(3 * (D // 2) ** 2 + 1)
    k = (2.0 * np.pi / float(box_len)) * np.sqrt(m.astype(np.float64))
    b = np.searchsorted(np.asarray(edges, dtype=np.float64), k, side="left").astype(np.int32)
    b[0] = -1
    return b


def power_spectrum(field, box_len, edges, field_sum=None, dimensional=False, lut=None):
    """(ps_sum, k_sum, counts) device tensors for the spherically binned
    Delta^2(k) of (field/mean - 1)."""
    D = field.shape[0]
    _check_box(field, D, "field")
    nb = len(edges)
    if field_sum is None:
        field_sum = field.double().sum().reshape(1)
    if lut is None:
        lut = torch.from_numpy(bin_lut(D, box_len, edges)).to(_DEV)
    if lut.dtype != torch.int32 or lut.numel() != 3 * (D // 2) ** 2 + 1 or not lut.is_cuda:
        raise ValueError(f"bin LUT must be a CUDA int32 tensor of {3 * (D // 2) ** 2 + 1} entries")
    if not 1 <= nb <= 512:
        raise ValueError("1 <= number of bins <= 512")
    ps = torch.empty(nb, dtype=torch.float64, device=_DEV)
    ks = torch.empty(nb, dtype=torch.float64, device=_DEV)
    cnt = torch.empty(nb, dtype=torch.int64, device=_DEV)
    nbytes = _lib.query("cm21_pk_workspace_bytes", D, nb)
    work = torch.empty(nbytes, dtype=torch.uint8, device=_DEV)
    _lib.call("cm21_power_spectrum", _lib.plan(D).handle, _lib.ptr(field), _lib.ptr(field_sum), D, float(box_len),
              _lib.ptr(lut), nb, 1 if dimensional else 0, _lib.ptr(ps), _lib.ptr(ks), _lib.ptr(cnt), _lib.ptr(work),
              nbytes, _lib.stream_handle())
    return ps, ks, cnt
