"""sgpu_compact_noalloc -- the compaction with its scratch in a fixed device
array of the library (no allocation) -- against sgpu_compact, through the
C-ABI: two calls in a row, the second on fewer packets than the first (the
array then still holds the first call's block counts past the second's),
each with a seeded mix of fates; the survivors' every field bit-identical
and the alive / escaped counts equal.  A count above the fixed limit is
rejected with SGPU_ERR_SIZE before anything runs, and the next call works."""
import ctypes

import numpy as np
import pytest

from pubsed_amd import abi, runtime

pytestmark = pytest.mark.gpu
FIELDS = ("x", "y", "z", "Dx", "Dy", "Dz", "t", "e", "nu", "ind", "type", "fate", "draw", "id")


def packets(n, seed):
    """A packet set with distinct values in every field and seeded fates."""
    rng = np.random.default_rng(seed)
    a = {k: rng.standard_normal(n) for k in ("x", "y", "z", "Dx", "Dy", "Dz", "t", "e", "nu")}
    a["ind"] = rng.integers(0, 100, n).astype(np.int32)
    a["type"] = rng.integers(0, 2, n).astype(np.int32)
    a["fate"] = rng.choice([abi.MOVING, abi.STOPPED, abi.ESCAPED, abi.ABSORBED], n,
                           p=[0.1, 0.4, 0.3, 0.2]).astype(np.int32)
    a["draw"] = rng.integers(0, 2 ** 32, n, dtype=np.uint64).astype(np.uint32)
    a["id"] = rng.integers(0, 2 ** 63, n, dtype=np.uint64)
    return a


def test_compact_noalloc_matches_compact(gpu):
    lib = gpu.lib
    cap = 700_000
    P = runtime.Problem(runtime.Space("cuda:0"), "c1_grey", nz=8)
    A = P.alloc_packets(cap, name="a")
    B = P.alloc_packets(cap, name="b")
    SA = P.alloc_packets(cap, name="sa")
    SB = P.alloc_packets(cap, name="sb")
    for n, seed in ((650_003, 1), (120_001, 2)):
        h = packets(n, seed)
        P.load_packets(h, name="a")
        P.load_packets(h, name="b")
        P.space.sync()
        na, ea, nb, eb = (ctypes.c_int64(-1) for _ in range(4))
        print("case compact n=%d" % n, flush=True)
        assert lib.sgpu_compact(gpu.ctx, ctypes.byref(A), n, ctypes.byref(SA), ctypes.byref(na),
                                ctypes.byref(ea)) == 0
        print("case compact_noalloc n=%d" % n, flush=True)
        assert lib.sgpu_compact_noalloc(gpu.ctx, ctypes.byref(B), n, ctypes.byref(SB), ctypes.byref(nb),
                                        ctypes.byref(eb)) == 0
        gpu.sync()
        keep = (h["fate"] != abi.ESCAPED) & (h["fate"] != abi.ABSORBED)
        assert na.value == nb.value == int(keep.sum())
        assert ea.value == eb.value == int((h["fate"] == abi.ESCAPED).sum())
        k = na.value
        ra, rb = P.packets_numpy(k, name="a"), P.packets_numpy(k, name="b")
        for f in FIELDS:
            assert np.array_equal(ra[f], rb[f]), f
            assert np.array_equal(rb[f], h[f][keep]), f


def test_compact_noalloc_rejects_more_than_its_limit(gpu):
    lib = gpu.lib
    P = runtime.Problem(runtime.Space("cuda:0"), "c1_grey", nz=8)
    A = P.alloc_packets(16, name="a")
    S = P.alloc_packets(16, name="s")
    big = (1 << 28) + 1
    # capacities claimed above the limit: the call must refuse before any launch
    A2 = abi.Packets(big, *[getattr(A, f) for f in FIELDS])
    S2 = abi.Packets(big, *[getattr(S, f) for f in FIELDS])
    na, ne = ctypes.c_int64(-1), ctypes.c_int64(-1)
    print("case compact_noalloc over the limit", flush=True)
    assert lib.sgpu_compact_noalloc(gpu.ctx, ctypes.byref(A2), big, ctypes.byref(S2), ctypes.byref(na),
                                    ctypes.byref(ne)) == -3
    buf = ctypes.create_string_buffer(512)
    lib.sgpu_last_error(buf, 512)
    assert b"limit" in buf.value
    h = packets(16, 3)
    P.load_packets(h, name="a")
    P.space.sync()
    assert lib.sgpu_compact_noalloc(gpu.ctx, ctypes.byref(A), 16, ctypes.byref(S), ctypes.byref(na),
                                    ctypes.byref(ne)) == 0
    gpu.sync()
    keep = (h["fate"] != abi.ESCAPED) & (h["fate"] != abi.ABSORBED)
    assert na.value == int(keep.sum())
