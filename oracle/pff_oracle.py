"""numpy front end of libpff_oracle.so (the C restatement in pff_oracle.c).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg.  Every function takes and returns numpy arrays.
"""

from __future__ import annotations

import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libpff_oracle.so")

c_int = ctypes.c_int32
c_i64 = ctypes.c_int64
c_d = ctypes.c_double
c_p = ctypes.c_void_p

MODES = {"normal": 0, "periodic_EW": 1, "periodic_NS": 2, "custom": 3}
INVALID_KEY = 0x4CBEBC200000002A


class OraGrid(ctypes.Structure):
    _fields_ = [("nx", c_int), ("ny", c_int), ("dx", ctypes.c_float), ("bound_mode", c_int), ("boundaries", c_p)]


class OraHydro(ctypes.Structure):
    _fields_ = [("prec", c_d), ("manning", c_d), ("edge_slope", c_d), ("dt", c_d), ("update", c_int)]


class OraLakeLog(ctypes.Structure):
    _fields_ = [
        ("max_log_iter", c_int),
        ("n_iter", c_int),
        ("n_loop", c_int),
        ("ndep0", c_i64),
        ("ndep", c_i64 * 64),
        ("bid_first", c_p),
        ("outlet_log", c_p),
        ("saddle_log", c_p),
        ("saddlenode_log", c_p),
    ]


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            from . import build as _b

            _b.build()
        _lib = ctypes.CDLL(LIB_PATH)
        _lib.ora_neighbour.restype = c_int
        _lib.ora_can_leave.restype = c_int
        _lib.ora_which_edge.restype = c_int
        _lib.ora_pack_key.restype = c_i64
        _lib.ora_pack_key.argtypes = [ctypes.c_float, c_int]
        _lib.ora_depression_counter.restype = c_i64
        _lib.ora_receivers_div_ambiguity.restype = c_i64
        _lib.ora_gradient_div_ambiguity.restype = c_i64
        _lib.ora_num_threads.restype = c_int
        _lib.ora_sentinel_hits.restype = c_i64
        _lib.ora_sentinel_hits.argtypes = [c_int]
        for name in ("ora_accumulate_constant_Q",):
            getattr(_lib, name).argtypes = [c_p, c_p, c_d, c_int, c_p]
        _lib.ora_fill_z_add_delta.argtypes = [c_p, c_p, c_p, c_p, c_d]
        _lib.ora_fill_z_add_delta_f64.argtypes = [c_p, c_p, c_p, c_p, c_d]
        _lib.ora_topofill.argtypes = [c_p, c_p, c_p, c_d]
        _lib.ora_pow53_array.argtypes = [ctypes.c_int64, c_p, c_p]
        _lib.ora_compute_receivers_stochastic.argtypes = [c_p, c_p, c_p, c_p, ctypes.c_uint64]
        _lib.ora_uniform01.argtypes = [ctypes.c_uint64, ctypes.c_int64, ctypes.c_int]
        _lib.ora_uniform01.restype = ctypes.c_float
    return _lib


def _p(a: np.ndarray):
    return a.ctypes.data_as(c_p)


class Grid:
    """Grid description; keeps the custom boundary array alive."""

    def __init__(self, nx, ny, dx=1.0, mode=0, boundaries=None):
        if isinstance(mode, str):
            mode = MODES[mode]
        self.nx, self.ny, self.dx, self.mode = int(nx), int(ny), float(dx), int(mode)
        self.b = None
        if boundaries is not None:
            self.b = np.ascontiguousarray(np.asarray(boundaries).ravel(), dtype=np.uint8)
        self.s = OraGrid(self.nx, self.ny, self.dx, self.mode, _p(self.b) if self.b is not None else None)
        self.N = self.nx * self.ny

    @property
    def ref(self):
        return ctypes.byref(self.s)


def _f32(a):
    return np.ascontiguousarray(a, dtype=np.float32).ravel()


def _i32(a):
    return np.ascontiguousarray(a, dtype=np.int32).ravel()


def neighbour(g: Grid, i: int, k: int) -> int:
    return lib().ora_neighbour(g.ref, int(i), int(k))


def can_leave(g: Grid, i: int) -> bool:
    return bool(lib().ora_can_leave(g.ref, int(i)))


def which_edge(g: Grid, i: int) -> int:
    return lib().ora_which_edge(g.ref, int(i))


def pack_key(f: float, i: int) -> int:
    return int(lib().ora_pack_key(float(np.float32(f)), int(i)))


def pack_array(f, idx):
    f = _f32(f)
    idx = _i32(idx)
    out = np.empty(f.size, np.int64)
    lib().ora_pack_array(c_i64(f.size), _p(f), _p(idx), _p(out))
    return out


def unpack_array(keys):
    keys = np.ascontiguousarray(keys, dtype=np.int64)
    f = np.empty(keys.size, np.float32)
    idx = np.empty(keys.size, np.int32)
    lib().ora_unpack_array(c_i64(keys.size), _p(keys), _p(f), _p(idx))
    return f, idx


def compute_receivers(g: Grid, z):
    z = _f32(z)
    rcv = np.empty(g.N, np.int32)
    grad = np.empty(g.N, np.float32)
    lib().ora_compute_receivers(g.ref, _p(z), _p(rcv), _p(grad))
    return rcv, grad


def compute_receivers_stochastic(g: Grid, z, seed: int):
    """receivers.py:15-58 with RAND_RCV: slopes weighted by seeded uniform
    draws (splitmix64 of seed + (4i+k) * golden ratio, 24 bits)."""
    z = _f32(z)
    rcv = np.empty(g.N, np.int32)
    grad = np.empty(g.N, np.float32)
    lib().ora_compute_receivers_stochastic(g.ref, _p(z), _p(rcv), _p(grad), ctypes.c_uint64(seed))
    return rcv, grad


def uniform01(seed: int, i: int, k: int) -> float:
    return float(lib().ora_uniform01(ctypes.c_uint64(seed), ctypes.c_int64(i), ctypes.c_int(k)))


def receivers_div_ambiguity(g: Grid, z) -> int:
    z = _f32(z)
    return int(lib().ora_receivers_div_ambiguity(g.ref, _p(z)))


def gradient_div_ambiguity(g: Grid, z) -> int:
    z = _f32(z)
    return int(lib().ora_gradient_div_ambiguity(g.ref, _p(z)))


def rcv2donor(g: Grid, rcv):
    rcv = _i32(rcv)
    dnr = np.zeros(4 * g.N, np.int32)
    ndnr = np.zeros(g.N, np.int32)
    lib().ora_rcv2donor(g.ref, _p(rcv), _p(dnr), _p(ndnr))
    return dnr, ndnr


def accumulate_constant_Q(g: Grid, rcv, value, area=True):
    """Reference schedule in f32; returns (Q, finished_within_cap)."""
    rcv = _i32(rcv)
    Q = np.empty(g.N, np.float32)
    done = lib().ora_accumulate_constant_Q(g.ref, _p(rcv), c_d(value), c_int(1 if area else 0), _p(Q))
    return Q, bool(done)


def accumulate_rake_compress(g: Grid, rcv, w, nrounds):
    rcv = _i32(rcv)
    w = _f32(w)
    Q = np.empty(g.N, np.float32)
    done = lib().ora_accumulate_rake_compress(g.ref, _p(rcv), _p(w), c_int(nrounds), _p(Q))
    return Q, bool(done)


def upstream_count(g: Grid, rcv):
    rcv = _i32(rcv)
    cnt = np.empty(g.N, np.uint32)
    rc = lib().ora_upstream_count(g.ref, _p(rcv), _p(cnt))
    if rc != 0:
        raise ValueError("receivers contain a cycle")
    return cnt


def accumulate_f64(g: Grid, rcv, w):
    rcv = _i32(rcv)
    w = np.ascontiguousarray(w, dtype=np.float64).ravel()
    Q = np.empty(g.N, np.float64)
    rc = lib().ora_accumulate_f64(g.ref, _p(rcv), _p(w), _p(Q))
    if rc != 0:
        raise ValueError("receivers contain a cycle")
    return Q


def depression_counter(g: Grid, rcv) -> int:
    rcv = _i32(rcv)
    return int(lib().ora_depression_counter(g.ref, _p(rcv)))


def basin_id_init(g: Grid):
    bid = np.empty(g.N, np.int32)
    lib().ora_basin_id_init(g.ref, _p(bid))
    return bid


def basin_identification(g: Grid, rec, npass=None):
    rec = _i32(rec)
    if npass is None:
        npass = int(np.ceil(np.log2(g.N))) + 1
    bid = np.empty(g.N, np.int32)
    lib().ora_basin_identification(g.ref, _p(rec), c_int(npass), _p(bid))
    return bid


def sentinel_hits(reset: bool = True) -> int:
    """Pass-5 removals (lakeflow.py:255-281) decided by reading an outlet that
    an earlier basin of the same pass had removed: the removed key is the
    sentinel pack(1e8, 42) and decodes to node 42."""
    return int(lib().ora_sentinel_hits(1 if reset else 0))


def saddlesort(g: Grid, bid, z, is_border=None):
    bid = _i32(bid)
    z = _f32(z)
    isb = np.zeros(g.N, np.uint8) if is_border is None else np.ascontiguousarray(is_border, np.uint8).copy()
    zp = np.empty(g.N, np.float32)
    bs = np.empty(g.N, np.int64)
    sn = np.empty(g.N, np.int32)
    ol = np.empty(g.N, np.int64)
    lib().ora_saddlesort(g.ref, _p(bid), _p(z), _p(isb), _p(zp), _p(bs), _p(sn), _p(ol))
    return dict(is_border=isb, z_prime=zp, basin_saddle=bs, saddlenode=sn, outlet=ol)


def reroute_jump(g: Grid, rec, outlet):
    rec = _i32(rec).copy()
    outlet = np.ascontiguousarray(outlet, np.int64)
    lib().ora_reroute_jump(g.ref, _p(rec), _p(outlet))
    return rec


def reroute_carve(g: Grid, rec_in, saddlenode, outlet):
    rec_in = _i32(rec_in)
    rec = rec_in.copy()
    sn = _i32(saddlenode)
    outlet = np.ascontiguousarray(outlet, np.int64)
    tag = np.zeros(g.N, np.uint8)
    it = lib().ora_reroute_carve(g.ref, _p(rec), _p(rec_in), _p(sn), _p(outlet), _p(tag))
    return rec, tag, int(it)


def reroute_flow(g: Grid, z, rec, carve=True, log_iters=0):
    """lakeflow.py:reroute_flow.  Returns dict with rec (router.receivers),
    rec_ (router.receivers_), bid_first and per-trip logs."""
    z = _f32(z)
    rec = _i32(rec).copy()
    rec_out = np.empty(g.N, np.int32)
    log = OraLakeLog()
    bid_first = np.zeros(g.N, np.int32)
    log.max_log_iter = log_iters
    log.bid_first = _p(bid_first)
    ol = sl = snl = None
    if log_iters > 0:
        ol = np.zeros((log_iters, g.N), np.int64)
        sl = np.zeros((log_iters, g.N), np.int64)
        snl = np.zeros((log_iters, g.N), np.int32)
        log.outlet_log = _p(ol)
        log.saddle_log = _p(sl)
        log.saddlenode_log = _p(snl)
    lib().ora_reroute_flow(g.ref, _p(z), _p(rec), _p(rec_out), c_int(1 if carve else 0), ctypes.byref(log))
    n = min(log.n_iter, log_iters)
    return dict(rec=rec, rec_=rec_out, bid_first=bid_first, ndep0=int(log.ndep0), n_iter=int(log.n_iter),
                n_loop=int(log.n_loop), ndep=[int(log.ndep[i]) for i in range(min(log.n_loop, 64))],
                outlet_log=None if ol is None else ol[:n], saddle_log=None if sl is None else sl[:n],
                saddlenode_log=None if snl is None else snl[:n])


def fill_z_add_delta(g: Grid, zh, h, rec, eps=1e-3):
    zh = _f32(zh).copy()
    h = _f32(h).copy()
    rec = _i32(rec)
    lib().ora_fill_z_add_delta(g.ref, _p(zh), _p(h), _p(rec), c_d(eps))
    return zh, h


def fill_z_add_delta_f64(g: Grid, zh, h, rec, eps=1e-3):
    zh = np.ascontiguousarray(zh, np.float64).ravel().copy()
    h = np.ascontiguousarray(h, np.float64).ravel().copy()
    rec = _i32(rec)
    lib().ora_fill_z_add_delta_f64(g.ref, _p(zh), _p(h), _p(rec), c_d(eps))
    return zh, h


def topofill(g: Grid, z, rec, eps=1e-4):
    z = _f32(z).copy()
    rec = _i32(rec)
    lib().ora_topofill(g.ref, _p(z), _p(rec), c_d(eps))
    return z


def pow53(h):
    """h^(5/3) rounded to f32, as the Manning update computes it (exact
    double operations; equal to f32(pow(double(h), 5/3)) on tested values)."""
    h = _f32(h)
    out = np.empty_like(h)
    lib().ora_pow53_array(h.size, _p(h), _p(out))
    return out


def hydro(prec=10e-3 / 3600, manning=0.033, edge_slope=1e-2, dt=5e-3, update=0):
    return OraHydro(prec, manning, edge_slope, dt, update)


def diffuse(g: Grid, hy, zh, Q, f64=False):
    if f64:
        zh = np.ascontiguousarray(zh, np.float64).ravel()
        Q = np.ascontiguousarray(Q, np.float64).ravel().copy()
        lib().ora_diffuse_Q_constant_prec_f64(g.ref, ctypes.byref(hy), _p(zh), _p(Q))
    else:
        zh = _f32(zh)
        Q = _f32(Q).copy()
        lib().ora_diffuse_Q_constant_prec(g.ref, ctypes.byref(hy), _p(zh), _p(Q))
    return Q


def manning(g: Grid, hy, h, zh, rec, Q, f64=False):
    rec = _i32(rec)
    if f64:
        h = np.ascontiguousarray(h, np.float64).ravel().copy()
        zh = np.ascontiguousarray(zh, np.float64).ravel().copy()
        Q = np.ascontiguousarray(Q, np.float64).ravel()
        lib().ora_graphflood_core_cte_mannings_f64(g.ref, ctypes.byref(hy), _p(h), _p(zh), _p(rec), _p(Q))
    else:
        h = _f32(h).copy()
        zh = _f32(zh).copy()
        Q = _f32(Q)
        lib().ora_graphflood_core_cte_mannings(g.ref, ctypes.byref(hy), _p(h), _p(zh), _p(rec), _p(Q))
    return h, zh


def graphflood_iteration(g: Grid, hy, zh, h, Q, f64=False):
    rec = np.empty(g.N, np.int32)
    if f64:
        zh = np.ascontiguousarray(zh, np.float64).ravel().copy()
        h = np.ascontiguousarray(h, np.float64).ravel().copy()
        Q = np.ascontiguousarray(Q, np.float64).ravel().copy()
        lib().ora_graphflood_iteration_f64(g.ref, ctypes.byref(hy), _p(zh), _p(h), _p(Q), _p(rec))
    else:
        zh = _f32(zh).copy()
        h = _f32(h).copy()
        Q = _f32(Q).copy()
        lib().ora_graphflood_iteration(g.ref, ctypes.byref(hy), _p(zh), _p(h), _p(Q), _p(rec))
    return zh, h, Q, rec


def num_threads() -> int:
    return int(lib().ora_num_threads())
