"""The packed (f32, i32) key of the oracle against the reference's own test
vectors (tests/golden/pack_unpack_cases.json, taken from
archive/test_pack_unpack.py) and an independent closed form."""

import json
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
CASES = json.load(open(os.path.join(HERE, "golden", "pack_unpack_cases.json")))


def closed_form(f, idx):
    """key = hi<<32 | idx with hi = bits(f) for f >= +0, else
    0x80000000 | (~bits(f) & 0x7fffffff), read as a signed 64-bit integer."""
    u = np.asarray(f, np.float32).view(np.uint32).astype(np.uint64)
    neg = (u & 0x80000000) != 0
    hi = np.where(neg, 0x80000000 | (~u & 0x7FFFFFFF), u) & 0xFFFFFFFF
    k = (hi << np.uint64(32)) | (np.asarray(idx, np.int64).astype(np.uint64) & 0xFFFFFFFF)
    return k.view(np.int64)


def test_round_trip_cases(ora):
    for name in ("round_trip", "simple"):
        f = np.array([c[0] for c in CASES[name]], np.float32)
        i = np.array([c[1] for c in CASES[name]], np.int32)
        keys = ora.pack_array(f, i)
        f2, i2 = ora.unpack_array(keys)
        np.testing.assert_array_equal(f2.view(np.uint32), f.view(np.uint32))
        np.testing.assert_array_equal(i2, i)


def test_orderings(ora):
    for name in ("lexicographic_order", "positive_order"):
        f = np.array([c[0] for c in CASES[name]], np.float32)
        i = np.array([c[1] for c in CASES[name]], np.int32)
        keys = ora.pack_array(f, i)
        assert np.all(keys[:-1] < keys[1:]), name


def test_massive_matches_tuple_sort(ora):
    m = CASES["massive"]
    np.random.seed(m["seed"])
    floats = np.random.uniform(m["float_range"][0], m["float_range"][1], m["n"]).astype(np.float32)
    idx = np.random.randint(0, m["index_high"], m["n"], dtype=np.int32)
    keys = ora.pack_array(floats, idx)
    by_key = [(floats[k], idx[k]) for k in np.argsort(keys, kind="stable")]
    assert by_key == sorted(zip(floats, idx))


def test_sentinel(ora):
    s = CASES["sentinel"]
    assert ora.pack_key(s["float"], s["index"]) == int(s["key_hex"], 16)


def test_closed_form_on_random_bits(ora):
    rng = np.random.default_rng(7)
    bits = rng.integers(0, 2**32, 20000, dtype=np.uint64).astype(np.uint32)
    f = bits.view(np.float32)
    f = np.where(np.isnan(f), np.float32(1.0), f).astype(np.float32)
    f[:4] = [0.0, -0.0, np.inf, -np.inf]
    idx = rng.integers(0, 2**31 - 1, 20000).astype(np.int32)
    np.testing.assert_array_equal(ora.pack_array(f, idx), closed_form(f, idx))


def test_negative_zero_orders_first(ora):
    assert ora.pack_key(-0.0, 5) < ora.pack_key(0.0, 5)
