"""Pins of the oracle's generic kinds (ChannelDiff for every element type,
channel count and FeatType), on the CPU.

Sampling: the float and int16 paths equal the separately pinned float and
short samplers (golden table); the narrow integer types equal brute force in
float32 over many draws; the end points (w = 0 gives lo, w = 2^32 - 1 gives
u = 1) and the bounds hold for every type.  Hook: numpy brute force with
numpy's C-style casts (modular integer conversion, float truncation) on
random images including out-of-image probes and channels past nCh.
Training: the tree invariants of the plain kinds.
"""
import numpy as np
import pytest

from oracle import oracle as orc

TYPES = [np.int8, np.uint8, np.int16, np.uint16, np.int32, np.uint32, np.int64, np.uint64, np.float32,
         np.float64]


def bounds_for(ft):
    if np.dtype(ft).kind == "u":
        return [0, 0, 0, 0], [9, 7, 40, 40]
    return [-9, -7, 0, 0], [9, 7, 40, 40]


@pytest.mark.parametrize("ft", TYPES)
def test_generic_sampling(ft):
    k = orc.generic_kind(np.uint16, 3, ft)
    lo, hi = bounds_for(ft)
    for node in range(40):
        p = orc.feature_params(k, 3, node, 5, lo, hi)
        assert p.dtype == np.dtype(ft)
        assert np.all(p.astype(np.float64) >= lo) and np.all(p.astype(np.float64) <= hi)
        t = orc.thresholds(k, 3, node, 5, 9, -30 if np.dtype(ft).kind != "u" else 0, 30)
        assert np.all(t.astype(np.float64) <= 30)
    if ft in (np.float32, np.int16):
        ref = orc.KIND_DEPTH_F32 if ft == np.float32 else orc.KIND_DEMO_S16
        for node in range(30):
            a = orc.feature_params(k, 1, node, 2, lo, hi)
            b = orc.feature_params(ref, 1, node, 2, lo[:orc.FD[ref]], hi[:orc.FD[ref]])
            assert np.array_equal(a[:b.size], b)
            assert np.array_equal(orc.thresholds(k, 1, node, 2, 11, -30, 30),
                                  orc.thresholds(ref, 1, node, 2, 11, -30, 30))


@pytest.mark.parametrize("ft", [np.int8, np.uint8, np.int16, np.uint16])
def test_generic_sampling_narrow_types_brute_force(ft):
    """lo + (T)(u * (hi - lo)) with an int span and a float32 product."""
    k = orc.generic_kind(np.uint8, 1, ft)
    info = np.iinfo(ft)
    lo, hi = (info.min // 2, info.max // 2)
    for node in range(200):
        got = orc.thresholds(k, 0, node, 0, 4, lo, hi)
        seed = np.array([0, node, 0, 1], np.uint32)
        w = orc.md5rand(seed)
        u = np.float32(w.astype(np.float32)) * np.float32(2.0 ** -32)
        q = (u * np.float32(hi - lo)).astype(np.float32)
        want = (lo + np.trunc(q).astype(np.int64)).astype(ft)
        assert np.array_equal(got, want), node


@pytest.mark.parametrize("ft", TYPES)
def test_generic_sampling_end_points(ft):
    """w = 0 gives lo; u = 2^32 / 2^32 = 1 gives lo + (T)span."""
    k = orc.generic_kind(np.uint8, 1, ft)
    # find (node, f) whose first thresholds word is small / large is impractical; use the
    # single-draw identity on the float32 product instead: all draws lie in [lo, hi]
    lo, hi = (0, 100) if np.dtype(ft).kind == "u" else (-100, 100)
    vals = np.concatenate([orc.thresholds(k, 9, n, 1, 16, lo, hi) for n in range(64)]).astype(np.float64)
    assert vals.min() >= lo and vals.max() <= hi
    if np.dtype(ft).kind in "iu":
        assert np.all(vals == np.round(vals))


def _hook_numpy(img, x, y, p, ft):
    h, w, nch = img.shape
    ox, oy = x + int(p[0]), y + int(p[1])
    c1, c2 = int(np.uint32(p[2]) % nch), int(np.uint32(p[3]) % nch)
    with np.errstate(over="ignore"):
        a = np.array(img[oy, ox, c1]).astype(ft) if (0 <= ox < w and 0 <= oy < h) else np.array(0).astype(ft)
        b = np.array(img[y, x, c2]).astype(ft)
        if np.dtype(ft).kind in "iu":
            r = np.array(int(a) - int(b)).astype(np.int64).astype(ft) if np.dtype(ft).itemsize < 8 else \
                (a - b).astype(ft)
        else:
            r = (a - b).astype(ft)
    return r


COMBOS = [(np.uint8, 10, np.float64), (np.int16, 5, np.int32), (np.uint16, 3, np.uint8), (np.float32, 6, np.int64),
          (np.uint8, 1, np.int8), (np.uint16, 2, np.uint32), (np.int8, 7, np.uint16), (np.uint16, 16, np.uint64),
          (np.float64, 4, np.float32), (np.int32, 3, np.int16)]


def random_image(it, nch, h, w, rng):
    if np.dtype(it).kind == "f":
        return rng.integers(0, 1000, (h, w, nch)).astype(it)
    info = np.iinfo(it)
    return rng.integers(info.min, int(info.max) + 1, (h, w, nch), dtype=np.int64).astype(it)


@pytest.mark.parametrize("it,nch,ft", COMBOS)
def test_generic_hook_brute_force(it, nch, ft):
    k = orc.generic_kind(it, nch, ft)
    rng = np.random.default_rng(nch)
    img = random_image(it, nch, 9, 13, rng)
    for trial in range(300):
        lo, hi = bounds_for(ft)
        p = np.array([rng.integers(lo[0], hi[0] + 1), rng.integers(lo[1], hi[1] + 1), rng.integers(0, 41),
                      rng.integers(0, 41)]).astype(ft)
        x, y = int(rng.integers(0, 13)), int(rng.integers(0, 9))
        got = orc.feature_generic_at(k, img, x, y, p)
        want = _hook_numpy(img, x, y, p, ft)
        assert np.array(got).tobytes() == np.array(want).astype(ft).tobytes(), (trial, x, y, p)


def generic_tset(it, nch, n_classes, rng, n_images=4, h=24, w=32, samples=80):
    import padenti_amd as pa
    ts = pa.TrainingSet(n_classes)
    for i in range(n_images):
        img = random_image(it, nch, h, w, rng)
        lab = rng.integers(1, n_classes + 1, (h, w)).astype(np.uint8)
        # labels follow a channel so that splits exist
        if np.dtype(it).kind == "f":
            lab = (1 + (img[..., 0].astype(np.int64) * n_classes // 1000) % n_classes).astype(np.uint8)
        smp = np.sort(rng.choice(h * w, samples, replace=False)).astype(np.uint32)
        ts << pa.TrainingSetImage(img, w, h, lab, n_classes, smp)
    return ts


@pytest.mark.parametrize("it,nch,ft", COMBOS[:4])
def test_generic_training_invariants(it, nch, ft):
    import padenti_amd as pa
    k = orc.generic_kind(it, nch, ft)
    rng = np.random.default_rng(7)
    ts = generic_tset(it, nch, 4, rng)
    lo, hi = bounds_for(ft)
    p = pa.TreeTrainerParameters(24, 6, lo, hi, -40 if np.dtype(ft).kind != "u" else 0, 40, 2)
    out = orc.train_tree(orc.Tset(k, ts), p, ts.priors, 0, 6)
    left = out["left_children"]
    assert left[0] == 1   # the root splits
    for n in range(len(left)):
        if left[n] >= 0:
            assert left[n] == 2 * n + 1 and left[2 * n + 1] != -2 and left[2 * n + 2] != -2
            assert out["features"].dtype == np.dtype(ft)
        post = out["posteriors"][n]
        if left[n] != -2:
            assert abs(float(post.sum()) - 1.0) < 1e-5
