"""TEST INFRASTRUCTURE -- numpy restatement of the 30-bit voxel codes.

Only ``tests/`` and ``bench.py``'s baseline leg use this module, as the
checker.  The product package never imports it.

Restated from src/serialize/z_order.cu (expandBits / extractBits and the
x*4 + y*2 + z interleave) and src/serialize/hilbert.cu (Skilling's
transform, J. Skilling, "Programming the Hilbert curve", AIP Conf. Proc. 707,
2004), on uint32 arrays with the reference's wrap-around arithmetic.
"""

from __future__ import annotations

from typing import Tuple

import numpy as np

U = np.uint32


def _u32(a) -> np.ndarray:
    """int32 / int64 values -> their low 32 bits as uint32 (torch's .int())."""
    return (np.asarray(a).astype(np.int64) & 0xFFFFFFFF).astype(U)


def spread3(v: np.ndarray) -> np.ndarray:
    v = _u32(v)
    with np.errstate(over="ignore"):
        v = (v * U(0x00010001)) & U(0xFF0000FF)
        v = (v * U(0x00000101)) & U(0x0F00F00F)
        v = (v * U(0x00000011)) & U(0xC30C30C3)
        v = (v * U(0x00000005)) & U(0x49249249)
    return v


def gather3(v: np.ndarray) -> np.ndarray:
    v = _u32(v) & U(0x49249249)
    v = (v ^ (v >> U(2))) & U(0x030C30C3)
    v = (v ^ (v >> U(4))) & U(0x0300F00F)
    v = (v ^ (v >> U(8))) & U(0x030000FF)
    v = (v ^ (v >> U(16))) & U(0x000003FF)
    return v


def _interleave(a, b, c) -> np.ndarray:
    with np.errstate(over="ignore"):
        return spread3(a) * U(4) + spread3(b) * U(2) + spread3(c)


def z_order_encode(x, y, z) -> np.ndarray:
    return _interleave(x, y, z).view(np.int32)


def z_order_decode(code) -> Tuple[np.ndarray, np.ndarray, np.ndarray]:
    c = _u32(code)
    return (gather3(c >> U(2)).view(np.int32), gather3(c >> U(1)).view(np.int32), gather3(c).view(np.int32))


def hilbert_encode(x, y, z) -> np.ndarray:
    p = [_u32(x).copy(), _u32(y).copy(), _u32(z).copy()]
    q = 1 << 9
    while q > 1:
        m = U(q - 1)
        for i in range(3):
            hit = (p[i] & U(q)) != 0
            t = (p[0] ^ p[i]) & m
            p0_hit = p[0] ^ m
            p0_miss = p[0] ^ t
            pi_miss = p[i] ^ t
            if i == 0:
                # p[i] is p[0]: the swap leaves it unchanged ((p0 ^ p0) & m = 0)
                p[0] = np.where(hit, p0_hit, p[0])
            else:
                p[i] = np.where(hit, p[i], pi_miss)
                p[0] = np.where(hit, p0_hit, p0_miss)
        q >>= 1
    p[1] = p[1] ^ p[0]
    p[2] = p[2] ^ p[1]
    t = np.zeros_like(p[2])
    q = 1 << 9
    while q > 1:
        t = np.where((p[2] & U(q)) != 0, t ^ U(q - 1), t)
        q >>= 1
    return _interleave(p[0] ^ t, p[1] ^ t, p[2] ^ t).view(np.int32)


def hilbert_decode(code) -> Tuple[np.ndarray, np.ndarray, np.ndarray]:
    c = _u32(code)
    p = [gather3(c >> U(2)), gather3(c >> U(1)), gather3(c)]
    t = p[2] >> U(1)
    p[2] = p[2] ^ p[1]
    p[1] = p[1] ^ p[0]
    p[0] = p[0] ^ t
    q = 2
    while q != (2 << 9):
        m = U(q - 1)
        for i in (2, 1, 0):
            hit = (p[i] & U(q)) != 0
            t = (p[0] ^ p[i]) & m
            if i == 0:
                p[0] = np.where(hit, p[0] ^ m, p[0])
            else:
                new0 = np.where(hit, p[0] ^ m, p[0] ^ t)
                p[i] = np.where(hit, p[i], p[i] ^ t)
                p[0] = new0
        q <<= 1
    return tuple(v.view(np.int32) for v in p)
