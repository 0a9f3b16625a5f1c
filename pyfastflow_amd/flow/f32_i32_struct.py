"""Order-preserving (f32, i32) -> i64 keys (pyfastflow/flow/f32_i32_struct.py).

As a signed 64-bit integer the key orders by the float (with -0.0 < +0.0),
then by the index, so one atomicMin finds a lexicographic argmin.
"""

from __future__ import annotations

import numpy as np
import torch

from .. import _lib
from ._core import check, ctx_of

__all__ = ["pack_full_float_index", "unpack_full_float_index", "pack_float_index", "unpack_float_index"]


def pack_float_index(f: float, i: int) -> int:
    """Host-side scalar key (f32_i32_struct.py:53-75)."""
    u = int(np.array([f], dtype=np.float32).view(np.uint32)[0])
    hi = (0x80000000 | (~u & 0x7FFFFFFF)) if (u & 0x80000000) else u
    v = (hi << 32) | (int(i) & 0xFFFFFFFF)
    return v - (1 << 64) if v >= (1 << 63) else v


def unpack_float_index(key: int):
    """Host-side inverse of pack_float_index (f32_i32_struct.py:78-103)."""
    v = int(key) & 0xFFFFFFFFFFFFFFFF
    hi = v >> 32
    u = (0x80000000 | (~hi & 0x7FFFFFFF)) if (hi & 0x80000000) else hi
    f = float(np.array([u], dtype=np.uint32).view(np.float32)[0])
    idx = int(np.array([v & 0xFFFFFFFF], dtype=np.uint32).view(np.int32)[0])
    return f, idx


def pack_full_float_index(arr, ft, it):
    """arr[i] = key(ft[i], it[i]) on the GPU (f32_i32_struct.py:123-135)."""
    ta = check("arr", arr, 0, torch.int64)
    n = ta.numel()
    tf = check("ft", ft, n, torch.float32)
    ti = check("it", it, n, torch.int32)
    _lib.call("pff_pack_full_float_index", ctx_of(ta), n, _lib.ptr(ta), _lib.ptr(tf), _lib.ptr(ti))


def unpack_full_float_index(arr, ft, it):
    """(ft[i], it[i]) = unpack(arr[i]) on the GPU (f32_i32_struct.py:105-121)."""
    ta = check("arr", arr, 0, torch.int64)
    n = ta.numel()
    tf = check("ft", ft, n, torch.float32)
    ti = check("it", it, n, torch.int32)
    _lib.call("pff_unpack_full_float_index", ctx_of(ta), n, _lib.ptr(ta), _lib.ptr(tf), _lib.ptr(ti))
