"""Field: the subset of Taichi's field API that PyFastFlow callers use, over a
1-D torch device tensor.

Callers of the reference use ``fill``, ``copy_from``, ``from_numpy`` and
``to_numpy`` on router/flooder fields (e.g. examples/exemple_raw_1.py:71-75,
examples/dev_graphflood.py:32-37), and ``field[None]`` on 0-d fields
(lakeflow.py:425-434).  Storage is a contiguous tensor in HBM; ``.data`` is
the tensor handed to the HIP library.
"""

from __future__ import annotations

import numpy as np
import torch

# Taichi dtype names used by the reference -> torch dtypes
f32 = torch.float32
f64 = torch.float64
i32 = torch.int32
i64 = torch.int64
u8 = torch.uint8
u1 = torch.uint8  # ti.u1 fields are stored one byte per element

_NP = {
    torch.float32: np.float32,
    torch.float64: np.float64,
    torch.int32: np.int32,
    torch.int64: np.int64,
    torch.uint8: np.uint8,
}


class Field:
    __slots__ = ("data",)

    def __init__(self, dtype=f32, shape=(), device=None, data: torch.Tensor | None = None):
        if data is None:
            if device is None:
                device = torch.device("cuda", torch.cuda.current_device())
            if isinstance(shape, int):
                shape = (shape,)
            data = torch.zeros(tuple(shape), dtype=dtype, device=device)
        self.data = data

    # --- Taichi field surface -------------------------------------------------
    @property
    def shape(self):
        return tuple(self.data.shape)

    @property
    def dtype(self):
        return self.data.dtype

    def fill(self, value):
        self.data.fill_(value)

    def copy_from(self, other):
        src = other.data if isinstance(other, Field) else other
        self.data.copy_(src)

    def from_numpy(self, arr):
        a = np.ascontiguousarray(np.asarray(arr).reshape(self.data.shape), dtype=_NP[self.data.dtype])
        self.data.copy_(torch.from_numpy(a))

    def to_numpy(self):
        return self.data.detach().cpu().numpy()

    def from_torch(self, t: torch.Tensor):
        self.data.copy_(t.reshape(self.data.shape))

    def to_torch(self) -> torch.Tensor:
        return self.data

    def __getitem__(self, key):
        if key is None:
            return self.data.item()
        return self.data[key].item()

    def __setitem__(self, key, value):
        if key is None:
            self.data.fill_(value)
        else:
            self.data[key] = value

    def __len__(self):
        return self.data.numel()

    def data_ptr(self):
        return self.data.data_ptr()

    def __repr__(self):
        return f"Field(shape={self.shape}, dtype={self.dtype}, device={self.data.device})"


def field(dtype=f32, shape=(), device=None) -> Field:
    """ti.field(dtype, shape=...) counterpart."""
    return Field(dtype=dtype, shape=shape, device=device)


def as_tensor(x) -> torch.Tensor:
    return x.data if isinstance(x, Field) else x
