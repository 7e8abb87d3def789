"""Pins the oracle's split search (A9/A10): deterministic log2, FP32 gain in
the reference op order (learn_best_feature.cl:42-93) and the threshold-major
first-maximum tie-break (:96-101 + std::max_element)."""
import math

import numpy as np

from oracle import oracle as orc
from tests import helpers as hp


def test_log2_det_within_one_ulp_and_exact_at_powers_of_two():
    rng = np.random.default_rng(3)
    xs = np.concatenate([rng.uniform(1e-6, 1.0, 20000), rng.uniform(0.5, 1.0, 5000),
                         np.float32(1.0) / np.arange(1, 3000, dtype=np.float32)]).astype(np.float32)
    for x in xs:
        got = orc.log2_det(x)
        ref = np.float32(math.log2(float(x)))
        ulp = np.spacing(np.float32(abs(ref))) if ref != 0 else np.float32(1e-45)
        assert abs(float(got) - float(ref)) <= float(ulp), (x, got, ref)
    for e in range(-20, 1):
        assert orc.log2_det(np.float32(2.0 ** e)) == np.float32(e)


def test_gain_f32_close_to_fp64_closed_form():
    rng = np.random.default_rng(4)
    for _ in range(3000):
        nC = int(rng.integers(2, 17))
        tot = rng.integers(0, 500, nC).astype(np.uint32)
        if tot.sum() == 0:
            tot[0] = 1
        left = np.array([rng.integers(0, t + 1) for t in tot], dtype=np.uint32)
        g32 = orc.gain_f32(left, tot)
        g64 = hp.gain64(left, tot)
        assert abs(float(g32) - g64) <= 2e-5 * max(1.0, math.log2(nC)), (left, tot, g32, g64)
        assert abs(orc.gain_f64(left, tot) - g64) <= 1e-12


def test_gain_special_cases():
    tot = np.array([10, 10, 0, 5], dtype=np.uint32)
    # all samples on one side: no information
    assert orc.gain_f32(np.zeros(4, np.uint32), tot) == np.float32(0.0)
    assert orc.gain_f32(tot, tot) == np.float32(0.0)
    # a perfect split of two equal pure classes gains one bit
    assert orc.gain_f32(np.array([8, 0], np.uint32), np.array([8, 8], np.uint32)) == np.float32(1.0)


def _hist_with(nC, nT, nF, best_pairs, rng):
    """H where every pair is a bad split except `best_pairs`, which are equal and pure."""
    tot = np.array([40] * nC, dtype=np.uint32)
    H = np.zeros((nC, nT, nF), dtype=np.uint32)
    for c in range(nC):
        H[c] = rng.integers(18, 23, size=(nT, nF))
    for (t, f) in best_pairs:
        H[:, t, f] = 0
        H[0, t, f] = 40
    return H, tot


def test_best_split_is_threshold_major_first_max():
    rng = np.random.default_rng(5)
    nC, nT, nF = 3, 2, 3
    # (t=0, f=2) is k = 2 in threshold-major order, (t=1, f=0) is k = 3;
    # feature-major order (f*nT + t) would rank them 4 and 1
    H, tot = _hist_with(nC, nT, nF, [(0, 2), (1, 0)], rng)
    k, g = orc.best_split(H, tot)
    assert k == 2
    assert g == orc.gain_f32(H[:, 0, 2], tot)


def test_best_split_matches_exhaustive_numpy():
    rng = np.random.default_rng(6)
    for _ in range(20):
        nC, nT, nF = int(rng.integers(2, 9)), int(rng.integers(1, 6)), int(rng.integers(1, 40))
        tot = rng.integers(1, 60, nC).astype(np.uint32)
        H = np.stack([rng.integers(0, t + 1, size=(nT, nF)) for t in tot]).astype(np.uint32)
        gains = np.array([[orc.gain_f32(H[:, t, f], tot) for f in range(nF)] for t in range(nT)],
                         dtype=np.float32).reshape(-1)
        k, g = orc.best_split(H, tot)
        assert k == int(np.argmax(gains)) and g == gains.max()   # argmax returns the first maximum
