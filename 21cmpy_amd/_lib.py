"""ctypes bindings of lib21cm_hip.so (declared in include/cm21.h).

torch is imported before the library is opened so that the library binds to
the HIP runtime and rocFFT that torch has already loaded (one runtime per
process).  There is no CPU fallback: every entry point raises when the library
is missing or returns an error status.
"""
import atexit
import ctypes
import os

import torch  # noqa: F401  (must precede the CDLL load, see module docstring)

LIB_PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), "lib21cm_hip.so")

CM21_OK = 0
STATUS_NAMES = {-1: "CM21_ERR_INVALID", -2: "CM21_ERR_HIP", -3: "CM21_ERR_FFT",
                -4: "CM21_ERR_DOMAIN", -5: "CM21_ERR_NOMEM"}

c_float_p = ctypes.POINTER(ctypes.c_float)
c_double_p = ctypes.POINTER(ctypes.c_double)


class CM21Error(RuntimeError):
    """A cm21_* call returned a non-zero status."""

    def __init__(self, func, status, message):
        self.func, self.status = func, status
        super().__init__(f"{func}: {STATUS_NAMES.get(status, status)}: {message}")


class BubbleParams(ctypes.Structure):
    """cm21_bubble_params (include/cm21.h)."""
    _fields_ = [
        ("box_len", ctypes.c_float),
        ("deltac", ctypes.c_float),
        ("zeta", ctypes.c_double),
        ("growth_inv", ctypes.c_float),
        ("fc_mean", ctypes.c_float),
        ("linear_input", ctypes.c_int),
        ("filter_type", ctypes.c_int),
        ("ref_fallthrough", ctypes.c_int),
        ("update", ctypes.c_int),
        ("sphere_periodic", ctypes.c_int),
        ("sphere_r_scale", ctypes.c_float),
        ("fused", ctypes.c_int),
        ("graph", ctypes.c_int),
    ]


class BatchBox(ctypes.Structure):
    """cm21_batch_box (include/cm21.h)."""
    _fields_ = [
        ("deltac", ctypes.c_float),
        ("growth_inv", ctypes.c_float),
        ("zeta", ctypes.c_double),
        ("fc_mean", ctypes.c_float),
        ("reserved", ctypes.c_int),
    ]


class BatchGroup(ctypes.Structure):
    """cm21_batch_group (include/cm21.h)."""
    _fields_ = [
        ("D", ctypes.c_int),
        ("nbox", ctypes.c_int),
        ("box_len", ctypes.c_float),
        ("n_radii", ctypes.c_int),
        ("radii", ctypes.POINTER(ctypes.c_float)),
        ("denoms", ctypes.POINTER(ctypes.c_float)),
        ("boxes", ctypes.POINTER(BatchBox)),
        ("delta", ctypes.c_void_p),
        ("ionized", ctypes.c_void_p),
        ("plan", ctypes.c_void_p),
        ("tile_columns", ctypes.c_int),
        ("tile_lines", ctypes.c_int),
    ]


class UnetConv(ctypes.Structure):
    """cm21_unet_conv (include/cm21.h)."""
    _fields_ = [
        ("B", ctypes.c_int), ("D", ctypes.c_int), ("cin", ctypes.c_int), ("cout", ctypes.c_int),
        ("c0", ctypes.c_int), ("o0", ctypes.c_int),
        ("x0", ctypes.c_void_p), ("x1", ctypes.c_void_p), ("mask", ctypes.c_void_p), ("w", ctypes.c_void_p),
        ("bias", ctypes.POINTER(ctypes.c_float)), ("bias_per_sample", ctypes.c_int), ("relu", ctypes.c_int),
        ("y0", ctypes.c_void_p), ("y1", ctypes.c_void_p), ("algo", ctypes.c_int),
        ("head", ctypes.POINTER(ctypes.c_float)), ("prob", ctypes.POINTER(ctypes.c_float)),
        ("pool", ctypes.c_void_p),
    ]


class UnetWgrad(ctypes.Structure):
    """cm21_unet_wgrad (include/cm21.h)."""
    _fields_ = [
        ("B", ctypes.c_int), ("D", ctypes.c_int), ("cin", ctypes.c_int), ("cout", ctypes.c_int), ("c0", ctypes.c_int),
        ("x0", ctypes.c_void_p), ("x1", ctypes.c_void_p), ("g", ctypes.c_void_p), ("mask", ctypes.c_void_p),
        ("dw", ctypes.POINTER(ctypes.c_float)), ("db", ctypes.POINTER(ctypes.c_float)),
        ("work", ctypes.c_void_p), ("work_bytes", ctypes.c_size_t),
    ]


VP, I, F, D, SZ, U32 = (ctypes.c_void_p, ctypes.c_int, ctypes.c_float, ctypes.c_double,
                         ctypes.c_size_t, ctypes.c_uint32)

# name -> (restype, argtypes)
SIGNATURES = {
    "cm21_abi_version": (I, []),
    "cm21_last_error": (ctypes.c_char_p, []),
    "cm21_plan_create": (I, [ctypes.POINTER(VP), I, I]),
    "cm21_plan_destroy": (I, [VP]),
    "cm21_r2c": (I, [VP, VP, VP, VP]),
    "cm21_c2r": (I, [VP, VP, VP, VP]),
    "cm21_hii_filter": (I, [VP, VP, I, F, I, I, I, F, F, VP]),
    "cm21_reduce_workspace_bytes": (SZ, []),
    "cm21_fcoll_sum": (I, [VP, I, F, F, VP, VP, VP]),
    "cm21_fcoll_threshold": (I, [VP, VP, I, F, F, D, VP, F, I, VP]),
    "cm21_sphere_workspace_bytes": (SZ, [I]),
    "cm21_update_sphere": (I, [VP, VP, I, F, F, D, VP, F, F, I, VP, VP]),
    "cm21_find_bubbles_workspace_bytes": (SZ, [I, I, I]),
    "cm21_find_bubbles": (I, [VP, VP, VP, c_float_p, c_float_p, I, ctypes.POINTER(BubbleParams), VP, SZ, VP]),
    "cm21_find_bubbles_noalloc": (I, [VP, VP, VP, c_float_p, c_float_p, I, ctypes.POINTER(BubbleParams), VP]),
    "cm21_grf_init": (I, [VP, I, D, D, D, D, U32, U32, VP]),
    "cm21_grf_init_table": (I, [VP, I, D, VP, VP, I, U32, U32, VP]),
    "cm21_subsample": (I, [VP, VP, I, I, F, VP]),
    "cm21_velocity": (I, [VP, VP, I, F, I, I, F, VP]),
    "cm21_move_mass": (I, [VP, VP, VP, VP, VP, I, I, F, F, I, VP]),
    "cm21_delta_T": (I, [VP, VP, VP, I, F, VP, VP, VP]),
    "cm21_pk_workspace_bytes": (SZ, [I, I]),
    "cm21_power_spectrum": (I, [VP, VP, VP, I, D, VP, I, I, VP, VP, VP, VP, SZ, VP]),
    "cm21_profile_enable": (I, [I]),
    "cm21_profile_collect": (I, [c_double_p, ctypes.POINTER(ctypes.c_int64), I]),
    "cm21_find_bubbles_batched_workspace_bytes": (SZ, [ctypes.POINTER(BatchGroup), I,
                                                       ctypes.POINTER(BubbleParams)]),
    "cm21_find_bubbles_batched": (I, [ctypes.POINTER(BatchGroup), I, ctypes.POINTER(BubbleParams), VP, SZ, VP]),
    "cm21_batch_tile_count": (I, [I, I]),
    "cm21_batch_tile_name": (ctypes.c_char_p, [I, I, I]),
    "cm21_batch_autotune": (I, [ctypes.POINTER(BatchGroup), I, ctypes.POINTER(BubbleParams), VP, SZ, I, c_double_p,
                                VP]),
    "cm21_unet_conv3_fwd": (I, [ctypes.POINTER(UnetConv), VP]),
    "cm21_unet_wgrad_workspace_bytes": (SZ, [I, I, I, I]),
    "cm21_unet_conv3_wgrad": (I, [ctypes.POINTER(UnetWgrad), VP]),
    "cm21_unet_pool_fwd": (I, [VP, VP, I, I, I, VP]),
    "cm21_unet_pool_bwd": (I, [VP, VP, I, I, I, I, VP]),
    "cm21_unet_pack_up": (I, [c_float_p, VP, VP, I, I, VP]),
    "cm21_unet_up_fwd": (I, [VP, VP, c_float_p, VP, I, I, I, I, VP]),
    "cm21_unet_up_workspace_bytes": (SZ, [I, I, I, I]),
    "cm21_unet_up_bwd": (I, [VP, VP, VP, VP, c_float_p, c_float_p, I, I, I, I, VP, SZ, VP]),
    "cm21_unet_in_fwd": (I, [c_float_p, c_float_p, c_float_p, c_float_p, c_float_p, VP, I, I, F, VP]),
    "cm21_unet_in_workspace_bytes": (SZ, [I, I]),
    "cm21_unet_in_wgrad": (I, [c_float_p, c_float_p, VP, VP, c_float_p, c_float_p, I, I, F, VP, SZ, VP]),
    "cm21_unet_head_workspace_bytes": (SZ, []),
    "cm21_unet_head": (I, [VP, c_float_p, c_float_p, c_float_p, c_float_p, VP, c_float_p, c_double_p, I, I, VP, SZ,
                           VP]),
    "cm21_unet_pack": (I, [c_float_p, VP, VP, I, I, VP]),
    "cm21_adamw": (I, [c_float_p, c_float_p, c_float_p, c_float_p, SZ, F, F, F, F, F, I, VP]),
}

STAGES = ["r2c", "filter", "c2r", "fcoll_sum", "threshold", "sphere", "fused_columns", "fused_rows",
          "fused_lines", "fused_paint", "fused_radius"]


def profile_enable(stages=True):
    """stages: True (all), False/None (off) or an iterable of stage names."""
    if stages is True:
        mask = -1
    elif not stages:
        mask = 0
    else:
        mask = 0
        for name in stages:
            mask |= 1 << STAGES.index(name)
    call("cm21_profile_enable", mask)


def profile_collect():
    """{stage: (total_ms, launches)} of the instrumented launches since the last collect."""
    n = len(STAGES)
    ms = (ctypes.c_double * n)()
    cnt = (ctypes.c_int64 * n)()
    call("cm21_profile_collect", ms, cnt, n)
    return {name: (ms[i], cnt[i]) for i, name in enumerate(STAGES) if cnt[i] > 0}

_lib = None


def load():
    """Open lib21cm_hip.so once and declare its signatures; raise if absent."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(f"{LIB_PATH} is not built; run `python 21cmpy_amd/build.py` "
                           "(there is no CPU fallback for the HIP path)")
    lib = ctypes.CDLL(LIB_PATH)
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    if lib.cm21_abi_version() != 1:
        raise RuntimeError(f"{LIB_PATH}: unexpected ABI version {lib.cm21_abi_version()}")
    _lib = lib
    return lib


def call(name, *args):
    """Call a status-returning entry point; raise CM21Error on failure."""
    lib = load()
    st = getattr(lib, name)(*args)
    if st != CM21_OK:
        msg = lib.cm21_last_error().decode(errors="replace")
        raise CM21Error(name, st, msg)
    return st


def query(name, *args):
    """Call a size-returning entry point."""
    return getattr(load(), name)(*args)


def ptr(t):
    """Device address of a CUDA tensor (checked)."""
    if t is None:
        return None
    if not t.is_cuda:
        raise ValueError("expected a device tensor")
    if not t.is_contiguous():
        raise ValueError("expected a contiguous tensor")
    return ctypes.c_void_p(t.data_ptr())


def stream_handle(stream=None):
    s = stream if stream is not None else torch.cuda.current_stream()
    return ctypes.c_void_p(s.cuda_stream)


def require_gpu():
    if not torch.cuda.is_available():
        raise RuntimeError("the cm21 HIP path needs a visible GPU (torch.cuda.is_available() is False)")


class Plan:
    """Owns a cm21_plan (rocFFT r2c/c2r plans for `batch` boxes of side D)."""

    def __init__(self, D, batch=1):
        require_gpu()
        h = ctypes.c_void_p()
        call("cm21_plan_create", ctypes.byref(h), int(D), int(batch))
        self.handle, self.D, self.batch = h, int(D), int(batch)

    def close(self):
        if self.handle:
            load().cm21_plan_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


_plans = {}


def plan(D, batch=1):
    """Cached plan per (D, batch) for the current process."""
    key = (int(D), int(batch))
    if key not in _plans:
        _plans[key] = Plan(D, batch)
    return _plans[key]


@atexit.register
def _release_plans():
    for p in _plans.values():
        try:
            p.close()
        except Exception:
            pass
    _plans.clear()
