"""21cmFAST box files: file-name parameters and raw float32 boxes.

Same interface and behaviour as the reference's tocmfastpy/boxio.py:
``parse_filename`` (:52-146) turns a box file name into a parameter dict,
``open_box`` (:209-253) reads raw float32 data, accepting the FFTW padded
layout (D, D, 2(D/2+1)), and ``readbox`` (:30-49) returns a Box with
``box_data`` and ``param_dict``.
"""
import os
import re

import numpy as np

_TYPE_PREFIXES = [
    ("updated_smoothed_deltax", "density"), ("xH_nohalos", "xh"), ("delta_T", "deltaT"),
    ("updated_vx", "vx"), ("updated_vy", "vy"), ("updated_vz", "vz"),
    ("smoothed_deltax", "density"), ("vxoverddot", "vx"), ("vyoverddot", "vy"), ("vzoverddot", "vz"),
    ("deltax", "init_x"), ("deltak", "init_k"),
]


def validate_box_type(prefix):
    """Box type from the leading letters-and-underscores of a file name (boxio.py:149-195)."""
    for tag, name in _TYPE_PREFIXES:
        if prefix.startswith(tag):
            return name
    return prefix


def parse_filename(filepath):
    base, filename = os.path.split(filepath)
    d = {}
    if not filename:
        d["type"] = False
        d["z"] = False
        return d
    d["filename"] = filename
    d["basedir"] = base
    m = re.search(r"([a-zA-Z]+_)+", filename)
    if m:
        d["type"] = validate_box_type(m.group())
    for key, pat, conv in (("nf", r"_nf([0-9.]+)", float), ("eff", r"_eff([0-9.]+)", float),
                           ("HIIfilter", r"_HIIfilter([0-9]+)", int), ("Mmin", r"_Mmin([0-9.e+-]+)", float),
                           ("RHIIfilter", r"_RHIImax([0-9.]+)", float)):
        m = re.search(pat, filename)
        if m:
            d[key] = conv(m.group(1))
    m = re.search(r"_z([0-9.]+)", filename)
    d["z"] = float(m.group(1).rstrip(".")) if m else 0.0
    m = re.search(r"_([0-9]+)_([0-9]+)Mpc", filename)
    if m:
        d["HIIdim"] = int(m.group(1))
        d["dim"] = int(m.group(1))
        d["BoxSize"] = int(m.group(2))
    m = re.search(r"_Iter([0-9]+)", filename)
    d["Iteration"] = int(m.group(1)) if m else 0
    m = re.search(r"_L([0-9]+)", filename)
    if m:
        d["BoxSize"] = int(m.group(1))
    return d


def trim_box(box):
    """Drop FFTW padding along the last axis (boxio.py:198-206)."""
    if box.shape[2] > box.shape[0]:
        return box[:, :, :box.shape[0]]
    return box


def open_box(filename, dim):
    data = np.fromfile(filename, dtype=np.float32)
    size = int(dim) ** 3
    if data.size == size:
        shape = (dim, dim, dim)
    else:
        root = int(round(data.size ** (1.0 / 3.0)))
        for r in (root - 1, root, root + 1):
            if r > 0 and r ** 3 == data.size:
                shape = (r, r, r)
                break
            if r > 0 and r * r * 2 * (r // 2 + 1) == data.size:
                shape = (r, r, 2 * (r // 2 + 1))
                break
        else:
            raise ValueError(f"{filename}: {data.size} floats is neither a cube nor an FFTW-padded cube")
    return trim_box(data.reshape(shape))


def save_box(filename, box_data):
    np.ascontiguousarray(box_data, dtype=np.float32).tofile(filename)


class Box:
    def __init__(self, box_data=None, param_dict=None):
        self.box_data = box_data
        self.param_dict = param_dict or {}

    def setBox(self, box_data, param_dict):
        self.box_data = box_data
        self.param_dict = param_dict


def readbox(filename, quiet=True):
    param_dict = parse_filename(filename)
    data = open_box(filename, param_dict["HIIdim"])
    return Box(data, param_dict)
