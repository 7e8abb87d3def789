"""Pins of the ingestion oracle (oracle/ingest.c) on the CPU.

CRC-32 against zlib and the check value; MT19937 against numpy's MT19937 under
legacy seeding and the C++ standard's requirement on the default seed;
Boost's uniform_int bucket against its closed form; both samplers on
hand-worked streams (rejection, ties of roundf, zero-size buckets, classes
above the bucket); the float arithmetic of the bucket sizes.
"""
import zlib

import numpy as np
import pytest

from oracle import ingest as oi


def test_crc32_check_value_and_zlib():
    assert oi.crc32(np.frombuffer(b"123456789", np.uint8)) == 0xCBF43926
    assert oi.crc32(np.zeros(0, np.uint8)) == 0
    rng = np.random.default_rng(3)
    for n in (1, 3, 63, 64, 65, 4095, 4096, 4097, 100003):
        d = rng.integers(0, 256, n, dtype=np.uint8)
        assert oi.crc32(d) == zlib.crc32(d.tobytes()), n
    u16 = rng.integers(0, 65536, (7, 5), dtype=np.uint16)   # bytes in host (little-endian) order
    assert oi.crc32(u16) == zlib.crc32(u16.tobytes())


def test_mt19937_standard_and_numpy():
    # [rand.predef]: the 10000th output of a default-constructed mt19937 is 4123659995
    assert oi.mt_outputs(5489, 10000)[-1] == 4123659995
    for seed in (0, 1, 12345, 0xDEADBEEF, 0xFFFFFFFF):
        ref = np.random.RandomState(seed)._bit_generator.random_raw(2000).astype(np.uint32)
        assert np.array_equal(oi.mt_outputs(seed, 2000), ref), seed


def _uniform_closed_form(stream, m, n):
    """floor(2^32 / m) buckets, rejection above m - 1."""
    bucket = (1 << 32) // m
    out, used = [], 0
    for x in stream:
        used += 1
        r = int(x) // bucket
        if r < m:
            out.append(r)
            if len(out) == n:
                break
    return out, used


@pytest.mark.parametrize("m", [2, 3, 4, 5, 7, 96, 1000, 65536, 76799, 307200])
def test_uniform_int_bucket_closed_form(m):
    w = m if m < 4096 else 640
    h = (m + w - 1) // w
    lab = np.zeros(w * h, np.uint8)
    lab[:m] = 1
    lab = lab.reshape(h, w)
    rng = np.random.default_rng(m)
    stream = rng.integers(0, 2 ** 32, 400, dtype=np.uint64).astype(np.uint32)
    # values just below / at the rejection edge of this bucket
    bucket = (1 << 32) // m
    edge = np.array([bucket * m - 1, bucket * m, 0xFFFFFFFF, bucket - 1, bucket], dtype=np.uint64)
    stream = np.concatenate([edge[edge < 2 ** 32].astype(np.uint32), stream])
    got, used = oi.sample_uniform_stream(lab, 64, stream)
    ranks, used2 = _uniform_closed_form(stream, m, 64)
    assert used == used2
    assert np.array_equal(got, np.sort(np.array(ranks, dtype=np.uint32)))   # pixel id == rank here


def test_uniform_hand_cases():
    lab = np.array([[0, 1, 0, 2], [3, 0, 0, 1]], np.uint8)   # labelled ids 1, 3, 4, 7
    s, used = oi.sample_uniform_stream(lab, 4, [0x00000000, 0xFFFFFFFF, 0x40000000, 0x7FFFFFFF])
    assert s.tolist() == [1, 3, 3, 7] and used == 4
    # m = 3: bucket 1431655765; only 0xFFFFFFFF rejects
    s, used = oi.sample_uniform_stream(np.ones((1, 3), np.uint8), 3,
                                       [0xFFFFFFFF, 1431655764, 0xFFFFFFFF, 2863311530, 1431655765])
    assert s.tolist() == [0, 1, 2] and used == 5
    # m = 2: bucket 2^31 (the +1 branch), 0xFFFFFFFF is accepted
    s, used = oi.sample_uniform_stream(np.array([[1, 0, 1]], np.uint8), 2, [0xFFFFFFFF, 0x7FFFFFFF])
    assert s.tolist() == [0, 2] and used == 2
    # one labelled pixel: no draw
    s, used = oi.sample_uniform_stream(np.array([[0, 0, 9]], np.uint8), 5, [])
    assert s.tolist() == [2] * 5 and used == 0
    # none: no sample
    s, used = oi.sample_uniform_stream(np.zeros((2, 2), np.uint8), 5, [])
    assert s.size == 0


def test_bucket_sizes_float_arithmetic():
    # 0.3f * 10 rounds to 3.0f (a tie to even); in double it would ceil to 4
    assert oi.bucket_sizes(10, 1, [0.3]).tolist() == [3]
    # 1.f / 3 * 100 = 33.333336f -> 34 each
    assert oi.bucket_sizes(100, 3).tolist() == [34, 34, 34]
    assert oi.bucket_sizes(2000, 2, [0.25, 0.0]).tolist() == [500, 0]


def test_weighted_hand_cases():
    lab = np.ones((1, 4), np.uint8)
    # B = 2: pixels 0, 1 fill; pixel 2 draws u = 0.9 -> roundf(1.8) = 2 (dropped);
    # pixel 3 draws u = 0.2 -> roundf(0.6) = 1 -> slot 1 = 3
    s, used = oi.sample_weighted_stream(lab, 2, 1, [0xE6666666, 0x33333333], weights=[1.0])
    assert s.tolist() == [0, 3] and used == 2
    # roundf rounds half away from zero: u = 0.25 at k = 2 -> 0.5 -> 1
    s, used = oi.sample_weighted_stream(lab, 2, 1, [0x40000000, 0x00000000], weights=[1.0])
    assert s.tolist() == [2, 3] and used == 2
    # a zero-size bucket still draws (class 2); class 1 keeps its last writer
    lab2 = np.array([[1, 2, 1, 2, 2, 1]], np.uint8)
    s, used = oi.sample_weighted_stream(lab2, 4, 2, [0xFFFFFFFF, 0x12345678, 0x9ABCDEF0, 0x80000000],
                                        weights=[0.5, 0.0])
    assert s.tolist() == [0, 5] and used == 4


def test_weighted_hand_case_unlabelled_pixels():
    # unlabelled pixels neither fill nor draw; B = 1
    lab3 = np.array([[0, 1, 0, 1, 0, 1]], np.uint8)
    s, used = oi.sample_weighted_stream(lab3, 1, 1, [0x00000000, 0xFFFFFFFF], weights=[1.0])
    # k = 1: u = 0 -> slot 0 = 3; k = 2: u ~ 1 -> roundf(2) = 2 >= 1 dropped
    assert s.tolist() == [3] and used == 2


def test_sampler_invariants_full_generator():
    rng = np.random.default_rng(5)
    for trial in range(6):
        h, w = rng.integers(1, 60, 2)
        nC = int(rng.integers(1, 9))
        lab = rng.integers(0, nC + 1, (h, w)).astype(np.uint8)
        lab[rng.random((h, w)) < 0.3] = 0
        data = rng.integers(0, 65535, (h, w), dtype=np.uint16)
        n = int(rng.integers(1, 300))
        s = oi.sample_uniform(data, lab, n, seed=trial)
        if (lab > 0).any():
            assert s.size == n and np.all(np.diff(s.astype(np.int64)) >= 0) and np.all(lab.reshape(-1)[s] > 0)
        else:
            assert s.size == 0
        wts = rng.random(nC).astype(np.float32)
        s2 = oi.sample_weighted(data, lab, n, nC, wts, seed=trial)
        B = oi.bucket_sizes(n, nC, wts)
        cnt = np.bincount(lab.reshape(-1), minlength=nC + 1)[1:]
        per = np.bincount(lab.reshape(-1)[s2], minlength=nC + 1)[1:]
        assert np.array_equal(per, np.minimum(cnt, B))
        assert np.all(np.diff(s2.astype(np.int64)) > 0)   # distinct pixels, sorted
        # classes that fit entirely keep every pixel
        flat = lab.reshape(-1)
        for c in range(nC):
            if cnt[c] <= B[c]:
                assert set(np.nonzero(flat == c + 1)[0]) <= set(s2.tolist())


def test_batch_equals_single():
    rng = np.random.default_rng(9)
    datas, labs = [], []
    for i in range(5):
        h, w = rng.integers(4, 40, 2)
        datas.append(rng.integers(0, 65535, (h, w), dtype=np.uint16))
        labs.append(rng.integers(0, 4, (h, w)).astype(np.uint8))
    res, crc = oi.sample_images(0, datas, labs, 50, seed=3)
    for d, l, r, c in zip(datas, labs, res, crc):
        assert np.array_equal(r, oi.sample_uniform(d, l, 50, seed=3))
        assert c == zlib.crc32(d.tobytes())
    res, _ = oi.sample_images(1, datas, labs, 20, n_classes=3, seed=3)
    for d, l, r in zip(datas, labs, res):
        assert np.array_equal(r, oi.sample_weighted(d, l, 20, 3, seed=3))


def test_rgb_to_labels_first_match():
    m = np.array([[255, 0, 0], [0, 0, 255], [255, 0, 0]], np.uint8)   # colour 0 repeated: first class wins
    px = np.array([[[255, 0, 0], [0, 0, 255], [1, 2, 3], [0, 0, 0]]], np.uint8)
    assert oi.rgb_to_labels(px, m).tolist() == [[1, 2, 0, 0]]
    px4 = np.concatenate([px, np.full((1, 4, 1), 7, np.uint8)], axis=2)   # alpha ignored
    assert oi.rgb_to_labels(px4, m).tolist() == [[1, 2, 0, 0]]
