"""Pins the oracle's sampler (A1-A3): MD5 keyed chain and the parameter /
threshold formula (hist_update.cl:62-184, prng.hpp:107-195)."""
import json
import os

import numpy as np
import pytest

from oracle import oracle as orc
from tests import helpers as hp

GOLD = os.path.join(os.path.dirname(__file__), "golden", "md5rand_vectors.json")


def _seeds(n, rng):
    s = [[0, 0, 0, 0], [0, 0, 0, 1], [1, 2, 3, 4], [0xffffffff] * 4, [7, 123456, 1999, 1]]
    s += rng.integers(0, 2 ** 32, size=(n, 4), dtype=np.uint64).tolist()
    return s


def test_md5rand_matches_rfc1321_md5():
    rng = np.random.default_rng(1)
    for seed in _seeds(300, rng):
        assert orc.md5rand(seed).tolist() == hp.md5rand(seed), seed


def test_md5rand_matches_reference_prng():
    ref = orc.ref_md5rand()
    if ref is None:
        pytest.skip("oracle/_ref not built (reference sources absent)")
    rng = np.random.default_rng(2)
    for seed in _seeds(300, rng):
        assert orc.md5rand(seed).tolist() == ref(seed).tolist(), seed


def test_md5rand_golden_vectors():
    with open(GOLD) as f:
        gold = json.load(f)
    assert len(gold["vectors"]) >= 16
    for v in gold["vectors"]:
        assert orc.md5rand(v["seed"]).tolist() == v["out"]


def test_md5_chain_words_reach_full_range():
    # (float)w / 0xFFFFFFFF can equal 1.0: w >= 0xFFFFFF80 rounds to 2^32
    assert hp.sample_f32(0xFFFFFFFF, -1.0, 1.0) == np.float32(1.0)
    assert float(orc.lib().or_sample_f32(0xFFFFFFFF, -1.0, 1.0)) == 1.0
    assert float(orc.lib().or_sample_f32(0, -60.0, 60.0)) == -60.0


@pytest.mark.parametrize("kind,fd,short", [(0, 4, False), (1, 4, False), (2, 6, False), (3, 2, True)])
def test_candidate_params_closed_form(kind, fd, short):
    lo = [-60.0, -60.0, -100.0, -100.0, 0.0, 0.0][:fd]
    hi = [60.0, 60.0, 100.0, 100.0, 4.0, 4.0][:fd]
    for tree, node, f in [(0, 0, 0), (0, 6, 99), (3, 1000, 1999), (2, 65534, 7), (0, 524287, 3999)]:
        got = orc.feature_params(kind, tree, node, f, lo, hi)
        want = hp.params(fd, tree, node, f, lo, hi, short)
        assert got.tobytes() == want.tobytes(), (tree, node, f)


@pytest.mark.parametrize("kind,short,lo,hi", [(0, False, -0.5, 0.5), (1, False, -500.0, 500.0),
                                              (3, True, -200, 200)])
def test_thresholds_closed_form(kind, short, lo, hi):
    for nT in (1, 3, 4, 5, 10, 20, 32):
        for tree, node, f in [(0, 0, 0), (1, 5, 77), (2, 40000, 1999)]:
            got = orc.thresholds(kind, tree, node, f, nT, lo, hi)
            want = hp.thresholds(nT, tree, node, f, lo, hi, short)
            assert got.tobytes() == want.tobytes()


def test_threshold_prefix_property():
    # threshold t depends only on md5 number t/4+1: longer lists extend shorter ones
    a = orc.thresholds(0, 0, 3, 5, 7, -1.0, 1.0)
    b = orc.thresholds(0, 0, 3, 5, 32, -1.0, 1.0)
    assert a.tobytes() == b[:7].tobytes()
