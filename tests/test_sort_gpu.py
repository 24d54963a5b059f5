"""The guiding build's own scan and stable radix sort (ap_guiding_scan_u32,
ap_guiding_sort_pairs_u64) against numpy.

Bar: bit-exact (integer work). The scan sizes cover the one-launch path (up
to 65,536 items), its edge and the three-launch path (block edges of 2048
items, a 1080p frame of paths); the sort sizes cover tile edges of 4096 keys
and partial bit ranges, with many equal keys so that stability shows.
"""
import numpy as np
import pytest

import aperture as ap

pytestmark = pytest.mark.gpu

SCAN_SIZES = [0, 1, 2047, 2048, 2049, 65_535, 65_536, 65_537, 131_072, 1_000_003, 1920 * 1080]


def test_scan_matches_numpy(gpu):
    g = np.random.default_rng(11)
    for n in SCAN_SIZES:
        a = g.integers(0, 6, n, dtype=np.uint32)
        if n > 5:
            a[n // 3] = 0xFFFF_FFF0  # a wrap in uint32 arithmetic
        want = np.zeros(n, dtype=np.uint64)
        if n:
            want[1:] = np.cumsum(a.astype(np.uint64))[:-1]
        want_total = int(a.astype(np.uint64).sum()) & 0xFFFF_FFFF
        got, total = ap.guiding_scan_u32(a)
        print(f"[sort] scan n {n}: total {total}", flush=True)
        assert np.array_equal(got, (want & 0xFFFF_FFFF).astype(np.uint32))
        assert total == want_total


@pytest.mark.parametrize("bits", [(0, 64), (0, 32), (5, 29)])
def test_radix_sort_is_stable_and_matches_numpy(gpu, bits):
    b, e = bits
    g = np.random.default_rng(12)
    for n in [1, 4095, 4096, 4097, 70_001, 300_007]:
        keys = g.integers(0, 1 << 62, n, dtype=np.uint64)
        keys[g.integers(0, n, n // 2)] = keys[g.integers(0, n, n // 2)]   # repeated keys
        keys[: n // 4] &= np.uint64(0xFFFF)                               # and many short ones
        vals = np.arange(n, dtype=np.uint32)
        mask = np.uint64((1 << (e - b)) - 1) if e - b < 64 else np.uint64(0xFFFF_FFFF_FFFF_FFFF)
        digits = (keys >> np.uint64(b)) & mask
        order = np.argsort(digits, kind="stable")
        got_k, got_v = ap.guiding_sort_pairs_u64(keys, vals, b, e)
        print(f"[sort] radix n {n} bits [{b}, {e})", flush=True)
        assert np.array_equal(got_v, vals[order])
        assert np.array_equal(got_k, keys[order])
