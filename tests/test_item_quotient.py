"""The (tracer, level) items of the bound kernel (kCf_node) split an item
index i into the tracer i / nw and the level i mod nw with a float product,
(int)((i + 0.5f) * (1.0f / nw)), instead of an integer division. The kernel
takes up to 4 tracers per block over up to 128 wet layers, so i < 4 nw; this
pins the identity exhaustively for i < 10 nw and nw <= 259, in IEEE binary32
round-to-nearest arithmetic, also with the reciprocal one ulp off either way
(a margin for a reciprocal that is not correctly rounded)."""
import numpy as np


def _quotients(nw, recip, n):
    i = np.arange(n, dtype=np.int64)
    q = (i.astype(np.float32) + np.float32(0.5)) * np.float32(recip)
    return np.floor(q.astype(np.float32)).astype(np.int64), i // nw


def test_float_item_quotient_is_exact():
    for nw in range(1, 260):
        r = np.float32(1.0) / np.float32(nw)
        for recip in (r, np.nextafter(r, np.float32(0.0)), np.nextafter(r, np.float32(1.0))):
            got, want = _quotients(nw, recip, 10 * nw)
            assert np.array_equal(got, want), f"nw={nw}, reciprocal {recip!r}"


def test_item_quotient_check_sees_a_wrong_rounding():
    """Without the half-item offset the truncated product misses exact
    multiples of nw for some nw, which the check above would report."""
    bad = 0
    for nw in range(1, 260):
        i = np.arange(10 * nw, dtype=np.int64)
        q = i.astype(np.float32) * (np.float32(1.0) / np.float32(nw))
        bad += int(np.sum(np.floor(q).astype(np.int64) != i // nw))
    assert bad > 0
