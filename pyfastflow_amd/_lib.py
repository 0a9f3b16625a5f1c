"""ctypes binding of libpff_hip.so (C ABI declared in include/pff.h).

The shared library must have been built in-tree (``python -m
pyfastflow_amd.build``).  There is no CPU fallback: if the library is missing
every compute call raises.
"""

from __future__ import annotations

import atexit
import ctypes
import os
import threading

import torch  # loaded first so its HIP runtime is the one the library binds to

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libpff_hip.so")

c_int = ctypes.c_int32
c_i64 = ctypes.c_int64
c_f = ctypes.c_float
c_d = ctypes.c_double
c_p = ctypes.c_void_p


class PffGrid(ctypes.Structure):
    _fields_ = [("nx", c_int), ("ny", c_int), ("dx", c_f), ("bound_mode", c_int), ("boundaries", c_p)]


class PffHydro(ctypes.Structure):
    _fields_ = [("prec", c_d), ("manning", c_d), ("edge_slope", c_d), ("dt", c_d), ("update", c_int)]


class PffLakeStatus(ctypes.Structure):
    _fields_ = [
        ("ndep0", c_i64),
        ("n_loop", c_int),
        ("n_iter", c_int),
        ("ndep", c_i64 * 64),
        ("carve_iters", c_int * 64),
    ]


class PffGfStatus(ctypes.Structure):
    _fields_ = [
        ("iterations", c_i64),
        ("residual", c_d),
        ("first_residual", c_d),
        ("converged", c_int),
        ("checks", c_int),
        ("nan_cells", c_i64),
        ("above_frac", c_d),
        ("first_above_frac", c_d),
        ("mean_change", c_d),
        ("first_mean_change", c_d),
        ("q50", c_d),
        ("q90", c_d),
        ("q99", c_d),
        ("q999", c_d),
        ("nan_iteration", c_i64),
    ]


class PffConvSrc(ctypes.Structure):
    _fields_ = [("mode", c_int), ("x", c_p), ("skip", c_p), ("planes", c_p), ("ca", c_int), ("nplanes", c_int),
                ("head_w", c_p), ("head_b", c_p), ("head_out", c_p)]


class PffTile(ctypes.Structure):
    _fields_ = [("nx", c_int), ("ny", c_int), ("z", c_p), ("rcv", c_p), ("Q", c_p)]


class PffBatchStatus(ctypes.Structure):
    _fields_ = [
        ("cells", c_i64),
        ("canvas_cells", c_i64),
        ("canvas_nx", c_int),
        ("canvas_ny", c_int),
        ("ndep0", c_i64),
        ("n_loop", c_int),
        ("per_tile", c_int),
        ("converged", c_int),
        ("capped_tiles", c_i64),
    ]


G = ctypes.POINTER(PffGrid)
H = ctypes.POINTER(PffHydro)

# name -> argtypes (every entry point returns int32 status unless listed in _RESTYPE)
SIGNATURES = {
    "pff_version": [],
    "pff_last_error": [],
    "pff_ctx_create": [ctypes.POINTER(c_p), c_int, c_p],
    "pff_ctx_destroy": [c_p],
    "pff_ctx_set_stream": [c_p, c_p],
    "pff_ctx_synchronize": [c_p],
    "pff_ctx_profile": [c_p, c_int],
    "pff_ctx_launch_overhead": [c_p, ctypes.c_int32, ctypes.POINTER(c_d)],
    "pff_ctx_profile_report": [c_p, ctypes.c_char_p, ctypes.POINTER(c_i64), ctypes.POINTER(c_d),
                               ctypes.POINTER(c_d)],
    "pff_compute_receivers": [c_p, G, c_p, c_p, c_p],
    "pff_compute_receivers_stochastic": [c_p, G, c_p, c_p, c_p, ctypes.c_uint64],
    "pff_rcv2donor": [c_p, G, c_p, c_p, c_p],
    "pff_rake_compress_accum": [c_p, G, c_p, c_p, c_p, c_p, c_p, c_p, c_p, c_int],
    "pff_fuse": [c_p, G, c_p, c_p, c_p, c_int],
    "pff_accumulate_constant_Q": [c_p, G, c_p, c_d, c_int, c_p, ctypes.POINTER(c_int)],
    "pff_accumulate_constant_Q_noalloc": [c_p, G, c_p, c_d, c_int, c_p, c_p],
    "pff_upstream_count": [c_p, G, c_p, c_p],
    "pff_depression_counter": [c_p, G, c_p, ctypes.POINTER(c_i64)],
    "pff_basin_id_init": [c_p, G, c_p],
    "pff_propagate_basin": [c_p, G, c_p, c_p],
    "pff_basin_identification": [c_p, G, c_p, c_p, c_p],
    "pff_saddlesort": [c_p, G, c_p, c_p, c_p, c_p, c_p, c_p, c_p],
    "pff_reroute_jump": [c_p, G, c_p, c_p],
    "pff_reroute_carve": [c_p, G, c_p, c_p, c_p, c_p, c_p, c_p, c_p],
    "pff_reroute_flow": [c_p, G, c_p, c_p, c_p, c_p, c_p, c_p, c_p, c_p, c_p, c_p, c_p, c_p, c_int, c_p,
                         ctypes.POINTER(PffLakeStatus)],
    "pff_ctx_lake_log": [c_p, c_p, c_p, c_p, c_i64, ctypes.POINTER(c_i64), c_int],
    "pff_batch_route": [c_p, c_int, ctypes.POINTER(PffTile), c_f, c_int, c_d, c_int,
                        ctypes.POINTER(PffBatchStatus)],
    "pff_conv_pack_weights": [c_p, c_int, c_int, c_p, c_p, c_p],
    "pff_conv3x3_fwd": [c_p, c_int, c_int, c_int, c_int, c_int, c_p, c_p, c_p, c_int, c_p],
    "pff_conv3x3_fwd_fused": [c_p, c_int, c_int, c_int, c_int, c_int, ctypes.POINTER(PffConvSrc), c_p, c_p, c_int,
                              c_p],
    "pff_conv3x3_bwd_data": [c_p, c_int, c_int, c_int, c_int, c_int, c_p, c_p, c_p],
    "pff_conv3x3_bwd_weight": [c_p, c_int, c_int, c_int, c_int, c_int, c_p, c_p, c_p, c_p],
    "pff_relu_bwd": [c_p, c_i64, c_p, c_p],
    "pff_avgpool2": [c_p, c_int, c_int, c_int, c_int, c_p, c_p],
    "pff_avgpool2_bwd": [c_p, c_int, c_int, c_int, c_int, c_p, c_p],
    "pff_upsample2": [c_p, c_int, c_int, c_int, c_int, c_p, c_p],
    "pff_upsample2_bwd": [c_p, c_int, c_int, c_int, c_int, c_p, c_p],
    "pff_concat": [c_p, c_i64, c_int, c_int, c_p, c_p, c_p],
    "pff_concat_bwd": [c_p, c_i64, c_int, c_int, c_p, c_p, c_p],
    "pff_head1x1": [c_p, c_i64, c_int, c_p, c_p, c_p, c_p],
    "pff_head1x1_bwd": [c_p, c_i64, c_int, c_p, c_p, c_p, c_p, c_p, c_p],
    "pff_pack_input": [c_p, c_i64, c_int, c_int, c_p, c_p],
    "pff_fill_z_add_delta": [c_p, G, c_p, c_p, c_p, c_p, c_p, c_p, c_d],
    "pff_topofill": [c_p, G, c_p, c_p, c_p, c_p, c_p, c_d],
    "pff_diffuse_Q_constant_prec": [c_p, G, H, c_p, c_p, c_p],
    "pff_graphflood_core_cte_mannings": [c_p, G, H, c_p, c_p, c_p, c_p, c_p],
    "pff_graphflood_run": [c_p, G, H, c_p, c_p, c_p, c_p, c_int, c_int, c_d, ctypes.POINTER(PffGfStatus)],
    "pff_graphflood_to_steady": [c_p, G, H, c_p, c_p, c_p, c_p, c_d, c_d, c_int, c_i64, c_int, c_d,
                                 ctypes.POINTER(PffGfStatus)],
    "pff_pack_full_float_index": [c_p, c_i64, c_p, c_p, c_p],
    "pff_unpack_full_float_index": [c_p, c_i64, c_p, c_p, c_p],
    "pff_fill_edges": [c_p, G, c_p],
    "pff_flow_out_nodes": [c_p, G, c_p],
    "pff_swap_arrays": [c_p, c_i64, c_p, c_p],
    "pff_init_arange": [c_p, c_i64, c_p],
}
_RESTYPE = {"pff_version": ctypes.c_char_p, "pff_last_error": ctypes.c_char_p}

_lib = None
_lock = threading.Lock()
_ctx = {}


def _destroy_contexts():
    """Release every context (its events, pinned words and scratch) while the
    HIP runtime is still up.  Registered after torch is imported, so Python
    runs it before torch's own exit hooks and before the runtime unloads."""
    lib = _lib
    if lib is None:
        return
    while _ctx:
        _, c = _ctx.popitem()
        try:
            lib.pff_ctx_destroy(c)
        except Exception:  # noqa: BLE001 - nothing useful to do at exit
            pass


atexit.register(_destroy_contexts)


def load():
    """Load libpff_hip.so once; raise if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    with _lock:
        if _lib is not None:
            return _lib
        if not os.path.exists(LIB_PATH):
            raise ImportError(
                f"{LIB_PATH} is missing: build it with `python -m pyfastflow_amd.build` "
                "(pyfastflow_amd has no CPU fallback)")
        lib = ctypes.CDLL(LIB_PATH)
        for name, args in SIGNATURES.items():
            fn = getattr(lib, name)
            fn.argtypes = args
            fn.restype = _RESTYPE.get(name, c_int)
        _lib = lib
    return _lib


class PffError(RuntimeError):
    pass


def call(name, *args):
    lib = load()
    rc = getattr(lib, name)(*args)
    if rc != 0:
        msg = lib.pff_last_error().decode(errors="replace")
        raise PffError(f"{name} failed ({rc}): {msg}")
    return rc


def ctx_for(device: torch.device | None = None):
    """Context bound to `device` and to torch's current stream on it."""
    if not torch.cuda.is_available():
        raise RuntimeError("pyfastflow_amd needs a ROCm GPU (torch.cuda.is_available() is False)")
    dev = torch.device("cuda", torch.cuda.current_device()) if device is None else torch.device(device)
    idx = dev.index if dev.index is not None else torch.cuda.current_device()
    load()
    c = _ctx.get(idx)
    stream = torch.cuda.current_stream(idx).cuda_stream
    if c is None:
        h = c_p()
        with torch.cuda.device(idx):
            call("pff_ctx_create", ctypes.byref(h), idx, c_p(stream))
        c = h
        _ctx[idx] = c
    else:
        call("pff_ctx_set_stream", c, c_p(stream))
    return c


def ptr(t) -> c_p:
    """Device address of a tensor or Field (None -> NULL)."""
    if t is None:
        return c_p(0)
    data = getattr(t, "data", t)
    return c_p(data.data_ptr())


def grid(nx: int, ny: int, dx: float, mode: int, boundaries=None) -> PffGrid:
    return PffGrid(int(nx), int(ny), float(dx), int(mode), ptr(boundaries) if boundaries is not None else c_p(0))


def hydro(prec: float, manning: float, edge_slope: float, dt: float, update: int = 0) -> PffHydro:
    return PffHydro(float(prec), float(manning), float(edge_slope), float(dt), int(update))


def declared_symbols(header: str | None = None) -> list[str]:
    """Entry points declared in include/pff.h (parsed from the header text)."""
    import re

    if header is None:
        header = os.path.join(_HERE, "..", "include", "pff.h")
    txt = open(header).read()
    return sorted(set(re.findall(r"\b(pff_[A-Za-z0-9_]+)\s*\(", txt)))
