"""Array utilities (pyfastflow/flow/util_taichi.py)."""

from __future__ import annotations

from .. import _lib
from ._core import check, ctx_of

__all__ = ["swap_arrays", "init_arange"]


def swap_arrays(array1, array2):
    """Swap two 4-byte-element arrays in place (util_taichi.py:18-35)."""
    a = check("array1", array1, 0)
    b = check("array2", array2, a.numel())
    if a.element_size() != 4 or b.element_size() != 4 or a.numel() != b.numel():
        raise ValueError("swap_arrays needs two arrays of equal length with 4-byte elements")
    _lib.call("pff_swap_arrays", ctx_of(a), a.numel(), _lib.ptr(a), _lib.ptr(b))


def init_arange(array):
    """array[i] = i (util_taichi.py:39-44)."""
    import torch

    a = check("array", array, 0, torch.int32)
    _lib.call("pff_init_arange", ctx_of(a), a.numel(), _lib.ptr(a))
