"""Pins the oracle's feature hooks (A5) against a numpy brute force:
out-of-image reads, zero depth, round half away from zero."""
import numpy as np
import pytest

from oracle import oracle as orc
from tests import helpers as hp


def _img(kind, rng, w=17, h=13):
    if kind == 0:
        im = rng.uniform(0.8, 5.0, (h, w)).astype(np.float32)
        im[rng.integers(0, h, 3), rng.integers(0, w, 3)] = 0.0
        im[2, 2] = 2.0
        return im
    if kind in (1, 3):
        im = rng.integers(500, 4000, (h, w)).astype(np.uint16)
        im[rng.integers(0, h, 3), rng.integers(0, w, 3)] = 0
        im[2, 2] = 2000
        if kind == 3:
            im[3, 3] = 40000   # wraps to a negative short
        return im
    im = rng.integers(0, 256, (h, w, 4)).astype(np.uint16)
    im[..., 3] = rng.integers(500, 4000, (h, w))
    im[2, 2, 3] = 2000
    return im


@pytest.mark.parametrize("kind", [0, 1, 2, 3])
def test_feature_brute_force(kind):
    rng = np.random.default_rng(10 + kind)
    img = _img(kind, rng)
    h, w = img.shape[0], img.shape[1]
    fd = orc.FD[kind]
    scale = {0: 30.0, 1: 30000.0, 2: 30000.0, 3: 40.0}[kind]
    for _ in range(2000):
        x, y = int(rng.integers(0, w)), int(rng.integers(0, h))
        p = rng.uniform(-scale, scale, fd)
        if kind == 2:
            p[4:] = rng.uniform(-0.5, 4.5, 2)
        p = p.astype(orc.FDT[kind])
        got = orc.feature_at(kind, img, x, y, p)
        want = hp.feature(kind, img, x, y, p)
        assert got == float(want), (x, y, p, got, want)


def test_round_half_away_from_zero_cases():
    img = np.full((9, 9), 2.0, dtype=np.float32)
    img[4, 4] = 2.0
    img[4, 7] = 7.0   # x + round(5/2 = 2.5) = 4 + 3 (half away); banker's rounding would give 6
    img[4, 6] = 6.0
    img[4, 1] = 1.0   # x + round(-5/2) = 4 - 3
    img[4, 2] = 3.0
    p = np.array([5.0, 0.0, -5.0, 0.0], dtype=np.float32)
    assert orc.feature_at(0, img, 4, 4, p) == 7.0 - 1.0


def test_out_of_image_reads_background():
    img = np.full((5, 5), 1000, dtype=np.uint16)
    p = np.array([1e6, 0.0, 0.0, 0.0], dtype=np.float32)   # far right -> 10000 mm
    assert orc.feature_at(1, img, 2, 2, p) == 10000.0 - 1000.0
    p = np.array([0.0, 0.0, 0.0, -1e9], dtype=np.float32)  # quotient beyond 2^20 -> outside
    assert orc.feature_at(1, img, 2, 2, p) == 1000.0 - 10000.0


def test_device_offset_rounding_equals_host_rule_exhaustively():
    """The device form of the offset rounding, trunc(q + copysign(0.5 - 2^-25, q))
    without a |q| < 0.5 case, equals the host rule (|q| < 0.5 -> 0, else
    trunc(q + copysign(0.5, q))) for every float32 in [0, 2^20]; the negative
    half follows by the sign symmetry of IEEE addition, copysign and trunc."""
    c = np.float32(0.49999997)
    assert c == np.nextafter(np.float32(0.5), np.float32(0))
    hi = int(np.float32(1048576.0).view(np.uint32))
    for start in range(0, hi + 1, 1 << 26):
        q = np.arange(start, min(start + (1 << 26), hi + 1), dtype=np.uint32).view(np.float32)
        host = np.where(q < np.float32(0.5), np.float32(0), np.trunc(q + np.float32(0.5)))
        dev = np.trunc(q + c)
        assert np.array_equal(host, dev), start
