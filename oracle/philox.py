"""Philox4x32-10 counter-based generator and Box-Muller normals in NumPy.

TEST INFRASTRUCTURE ONLY (see ``oracle/__init__.py``).

The reference draws its Gaussian random field with PyCUDA's MRG32k3a seeded
from Python's ``random`` (Programs/initialize.py:42,70-71), so its stream is
not reproducible.  The GPU path uses Philox4x32-10 (Salmon, Moraes, Dror and
Shaw, "Parallel random numbers: as easy as 1, 2, 3", SC'11), restated here
from the published algorithm: ten rounds of two 32x32->64 multiplies with the
round keys bumped by the Weyl constants.
"""
import numpy as np

PHILOX_M0 = np.uint64(0xD2511F53)
PHILOX_M1 = np.uint64(0xCD9E8D57)
PHILOX_W0 = np.uint32(0x9E3779B9)
PHILOX_W1 = np.uint32(0xBB67AE85)
_MASK32 = np.uint64(0xFFFFFFFF)


def philox4x32_10(ctr, key):
    """Vectorised Philox4x32-10.

    ctr: uint32 array of shape (..., 4); key: uint32 array broadcastable to (..., 2).
    Returns uint32 array (..., 4).
    """
    c = np.asarray(ctr, dtype=np.uint32)
    k = np.asarray(key, dtype=np.uint32)
    c0, c1, c2, c3 = (c[..., i].astype(np.uint64) for i in range(4))
    k0 = np.broadcast_to(k[..., 0], c0.shape).astype(np.uint32)
    k1 = np.broadcast_to(k[..., 1], c0.shape).astype(np.uint32)
    with np.errstate(over="ignore"):
        for rnd in range(10):
            if rnd > 0:
                k0 = (k0 + PHILOX_W0).astype(np.uint32)
                k1 = (k1 + PHILOX_W1).astype(np.uint32)
            p0 = PHILOX_M0 * c0
            p1 = PHILOX_M1 * c2
            hi0, lo0 = p0 >> np.uint64(32), p0 & _MASK32
            hi1, lo1 = p1 >> np.uint64(32), p1 & _MASK32
            n0 = hi1 ^ c1 ^ k0.astype(np.uint64)
            n1 = lo1
            n2 = hi0 ^ c3 ^ k1.astype(np.uint64)
            n3 = lo0
            c0, c1, c2, c3 = n0, n1, n2, n3
    return np.stack([c0, c1, c2, c3], axis=-1).astype(np.uint32)


def uniform_open01(u32):
    """Map uint32 to a float in (0,1): ((u >> 8) + 0.5) * 2^-24 (exact in fp32 and fp64)."""
    return ((np.asarray(u32, dtype=np.uint32) >> np.uint32(8)).astype(np.float64) + 0.5) * (2.0 ** -24)


def box_muller_f64(u0, u1):
    """Two standard normals from two uint32 words, computed in FP64.

    r = sqrt(-2 ln a), theta = 2 pi b, g1 = r cos(theta), g2 = r sin(theta).
    """
    a = uniform_open01(u0)
    b = uniform_open01(u1)
    r = np.sqrt(-2.0 * np.log(a))
    th = 2.0 * np.pi * b
    return r * np.cos(th), r * np.sin(th)
