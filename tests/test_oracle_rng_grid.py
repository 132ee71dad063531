"""Pins of the oracle's RNG and binned-grid arithmetic (CPU only).

* Philox-4x32-10 against the published Random123 known-answer vectors.
* locate_array::log_init against a replay of the repeated multiplication.
* locate / locate_within_bounds against brute-force searchsorted.
* cdf_array::normalize / sample semantics.
"""
import ctypes

import numpy as np
import pytest

from pubsed_amd import abi, workloads as W
from oracle import oracle_py

KAT = [  # Random123 kat_vectors, philox4x32 with 10 rounds
    ([0, 0, 0, 0], [0, 0], [0x6627e8d5, 0xe169c58d, 0xbc57ac4c, 0x9b00dbd8]),
    ([0xffffffff] * 4, [0xffffffff] * 2, [0x408f276d, 0x41c83b0e, 0xa20bc7c6, 0x6d5451fd]),
    ([0x243f6a88, 0x85a308d3, 0x13198a2e, 0x03707344], [0xa4093822, 0x299f31d0],
     [0xd16cfe09, 0x94fdcceb, 0x5001e420, 0x24126ea1]),
]


@pytest.mark.parametrize("ctr,key,out", KAT)
def test_philox_known_answers(ctr, key, out):
    lib = oracle_py.load()
    c, k, o = np.array(ctr, np.uint32), np.array(key, np.uint32), np.zeros(4, np.uint32)
    lib.or_philox4x32(c.ctypes.data, k.ctypes.data, o.ctypes.data)
    assert list(o) == out


def test_uniform_mapping_and_moments():
    lib = oracle_py.load()
    u = np.array([lib.or_uniform(99, 5, d) for d in range(20000)])
    assert u.min() >= 0.0 and u.max() < 1.0
    assert abs(u.mean() - 0.5) < 5 * np.sqrt(1 / 12 / len(u))
    # the double is the top 53 bits of words (w1:w0) of block {draw, 0, id}
    c = np.array([3, 0, 5, 0], np.uint32)
    k = np.array([99, 0], np.uint32)
    o = np.zeros(4, np.uint32)
    lib.or_philox4x32(c.ctypes.data, k.ctypes.data, o.ctypes.data)
    v = ((int(o[1]) << 32) | int(o[0])) >> 11
    assert lib.or_uniform(99, 5, 3) == v * 2.0 ** -53
    f = lib.or_uniform_f32(99, 5, 3)
    assert f == np.float32((int(o[0]) >> 8) * 2.0 ** -24)


def test_log_init_replay_and_counts():
    lib = oracle_py.load()
    x = np.zeros(5000)
    mn = ctypes.c_double()
    for lo, hi, n in [(1e14, 1e16, 1024), (W.C_LIGHT / 1e-4, W.C_LIGHT / 1e-5, 4096)]:
        g = W.nu_log_grid(lo, hi, n)
        m = lib.or_log_init(lo, hi, g["del_"], x.ctypes.data, 5000, ctypes.byref(mn))
        assert m == n
        assert np.array_equal(x[:m], g["x"])          # bit-identical edges
        assert mn.value == lo and x[m - 1] == hi
    # the exact (max/min)^(1/n)-1 leaves a sliver bin on 1e14..1e16 / 1024
    d = 100.0 ** (1 / 1024) - 1
    assert len(W.log_init(1e14, 1e16, d)) == 1025


def _loc(g):
    x = np.ascontiguousarray(g["x"])
    return abi.Locate(g["n"], g["type"], g["min"], g["del_"], x.ctypes.data), x


@pytest.mark.parametrize("kind", ["log", "log_edges", "lin", "flex"])
def test_locate_brute_force(kind):
    """locate against upper_bound (searchsorted right) on random values, on
    every edge and at min.  On the 4096-bin line grid the log formula lands
    one bin low at 4095 of its 4096 edges (the repeated-multiplication edges
    sit just below min (1+del)^k), so the verify-and-fall-back branch
    (locate_array.cpp:191-230) is taken at every one of them -- except at
    the first edge, where the formula's 0 is returned unverified (the
    reference verifies only 0 < ind < size): locate(x[0]) = 0 there, one
    below upper_bound.  On the 1024-bin grid the formula is exact at every
    edge."""
    lib = oracle_py.load()
    rng = np.random.default_rng(1)
    if kind == "log":
        g = W.nu_log_grid(1e14, 1e16, 1024)
    elif kind == "log_edges":
        g = W.nu_log_grid(W.C_LIGHT / 1e-4, W.C_LIGHT / 1e-5, 4096)
        misses = sum(int(np.floor(np.log(v / g["min"]) / np.log(1 + g["del_"])))
                     != np.searchsorted(g["x"], v, side="right") for v in g["x"][:-1])
        assert misses == 4095
    elif kind == "lin":
        g = W.lin_init_del(-3.0, 7.0, 0.37)
    else:
        g = dict(n=50, type=abi.LOC_FLEX, min=0.0, del_=0.0, x=np.sort(rng.uniform(0.1, 5, 50)))
    L, keep = _loc(g)
    lo, hi = g["min"], g["x"][-1]
    vals = np.concatenate([rng.uniform(lo - 0.1 * abs(hi), hi * 1.1, 4000), g["x"], [lo]])
    for v in vals:
        want = 0 if v < lo else int(np.searchsorted(g["x"], v, side="right"))
        if kind == "log_edges" and v == g["x"][0]:
            want = 0
        assert lib.or_locate(ctypes.byref(L), float(v)) == want
        assert lib.or_locate_within_bounds(ctypes.byref(L), float(v)) == min(want, g["n"] - 1)


def test_cdf_normalise_and_sample_semantics():
    """cdf_array: NaN entries become 0, an all-zero CDF becomes flat ones,
    sample() is upper_bound (first entry larger than the draw)."""
    lib = oracle_py.load()
    lib.or_cdf_normalize.argtypes = [ctypes.c_void_p, ctypes.c_int]
    lib.or_cdf_sample.restype = ctypes.c_int
    lib.or_cdf_sample.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_double]
    y = np.array([0.0, 1.0, 1.0, 3.0, 4.0])
    lib.or_cdf_normalize(y.ctypes.data, 5)
    assert np.array_equal(y, [0, 0.25, 0.25, 0.75, 1.0])
    assert lib.or_cdf_sample(y.ctypes.data, 5, 0.0) == 1
    assert lib.or_cdf_sample(y.ctypes.data, 5, 0.25) == 3
    assert lib.or_cdf_sample(y.ctypes.data, 5, 0.999) == 4
    z = np.zeros(4)
    lib.or_cdf_normalize(z.ctypes.data, 4)
    assert np.array_equal(z, np.ones(4))
    w = np.array([1.0, np.nan, 2.0])
    lib.or_cdf_normalize(w.ctypes.data, 3)
    assert np.array_equal(w, [0.5, 0.0, 1.0])
