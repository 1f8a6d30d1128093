"""CPU test of the k-buffer insertion the forward kernel uses.

The reference inserts a hit h into its sorted k-buffer with a swap chain
(src/GS_shader.ts:349-353): for each slot, if h precedes the slot's entry
the two swap.  The kernel keeps one 64-bit key per slot (t bits above the
Gaussian id below, so unsigned order is (t, id) order) and inserts with
    slot'[i] = min(slot[i], max(slot[i-1], h)),   slot'[0] = min(slot[0], h)
evaluated by the FP64 max/min instructions on the keys read as doubles.
This test checks, on random sorted buffers with dummy slots (key 0),
sentinels of t_max (finite or +inf), denormal and huge t and repeated keys,
that the min/max network on the double view equals the swap chain on the
integer keys, and that every key read as a double is a non-negative
non-NaN value.
"""
import numpy as np

SENTINEL_ID = 0xFFFFFFFF


def key(t, gid):
    return (int(np.float32(t).view(np.uint32)) << 32) | int(gid)


def swap_chain(buf, h):
    """GS_shader.ts:349-353 with the (t, id) key: strict compare, so equal
    keys keep their arrival order."""
    buf = list(buf)
    for i in range(len(buf)):
        if h < buf[i]:
            buf[i], h = h, buf[i]
    return buf


def min_max_network(buf, h):
    """The kernel's insertion: from the last slot down, in place, on the
    keys read as float64 (numpy's maximum/minimum on non-NaN doubles)."""
    b = np.array(buf, np.uint64).view(np.float64).copy()
    hd = np.array([h], np.uint64).view(np.float64)[0]
    assert not np.isnan(b).any() and not np.isnan(hd) and (b >= 0).all() and hd >= 0
    if not (h < buf[-1]):
        return list(buf)
    for i in range(len(b) - 1, 0, -1):
        b[i] = np.minimum(b[i], np.maximum(b[i - 1], hd))
    b[0] = np.minimum(b[0], hd)
    return [int(x) for x in b.view(np.uint64)]


def test_min_max_network_equals_swap_chain():
    rng = np.random.default_rng(7)
    ts = np.float32([1e-40, 1e-38, 1e-3, 0.25, 0.5, 1.0, 2.5, 7.0, 1e5, 3e38])
    cases = 0
    for _ in range(4000):
        maxk = int(rng.choice([4, 8, 16, 32]))
        k = int(rng.integers(1, maxk + 1))
        t_max = np.float32(rng.choice([1e5, 3e38, np.inf]))
        pool = ts[ts < t_max]
        # a buffer as the kernel holds it: MAXK - k dummies, then k active
        # slots filled with hits in order by arrival, sentinels beyond
        buf = [0] * (maxk - k) + [key(t_max, SENTINEL_ID)] * k
        for _ in range(int(rng.integers(0, 3 * k + 2))):
            h = key(rng.choice(pool), rng.integers(0, 6))
            want = swap_chain(buf, h)
            got = min_max_network(buf, h)
            assert got == want, (buf, h)
            assert all(got[i] <= got[i + 1] for i in range(maxk - 1))
            buf = got
            cases += 1
    assert cases > 20000


def test_double_view_orders_like_the_keys():
    """Non-negative, non-NaN doubles order like their bit patterns, and a
    key of a non-NaN t >= +0 (or +inf) is such a double."""
    rng = np.random.default_rng(3)
    ts = np.concatenate([np.float32([0.0, 1e-45, 1e-40, 1e-38, 3e38, np.inf]),
                         rng.uniform(0, 100, 500).astype(np.float32)])
    keys = np.array(sorted({key(t, i) for t in ts for i in (0, 1, 7, SENTINEL_ID)}), np.uint64)
    d = keys.view(np.float64)
    assert not np.isnan(d).any() and np.isfinite(d).all() and (d >= 0).all()
    assert (np.diff(d) > 0).all()  # strictly increasing with the integer keys
