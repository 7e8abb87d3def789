"""Pins the oracle's histogram (A6/A7), training loop (A10/A11/A13) and
prediction (A14/A15) by brute force and invariants."""
import numpy as np
import pytest

from oracle import oracle as orc
from padenti_amd import TrainingSet, TrainingSetImage, Tree, TreeTrainerParameters, synth
from tests import helpers as hp


def tiny_set(kind, rng, n_img=2, w=16, h=12, nC=4, ns=32):
    ts = TrainingSet(nC)
    for i in range(n_img):
        if kind == 0:
            px = rng.uniform(0.8, 5.0, (h, w)).astype(np.float32)
        else:
            px = rng.integers(500, 4000, (h, w)).astype(np.uint16)
        lab = rng.integers(1, nC + 1, (h, w)).astype(np.uint8)
        lab[0, :3] = 0   # a few unlabelled pixels (never sampled)
        cand = np.flatnonzero(lab.reshape(-1))
        smp = np.sort(rng.choice(cand, ns, replace=True)).astype(np.uint32)
        ts << TrainingSetImage(px, w, h, lab, nC, smp)
    return ts


@pytest.mark.parametrize("kind", [0, 1, 3])
def test_hist_brute_force(kind):
    rng = np.random.default_rng(20 + kind)
    ts = tiny_set(kind, rng)
    fd = orc.FD[kind]
    short = kind == 3
    b = {0: 30.0, 1: 30000.0, 3: 40}[kind]
    tb = {0: 2.0, 1: 2000.0, 3: 300}[kind]
    p = TreeTrainerParameters(8, 4, [-b] * fd, [b] * fd, -tb, tb, 0)
    o = orc.Tset(kind, ts)
    S = o.total
    nodes = [3, 0, 9]
    sample_node = rng.choice([0, 3, 9, -1], S).astype(np.int32)
    H = orc.hist_nodes(o, p, 7, nodes, sample_node)
    want = np.zeros_like(H)
    g = 0
    for im in ts.images:
        img = im.data.reshape(im.height, im.width)
        for s in im.samples:
            n = sample_node[g]
            g += 1
            if n < 0:
                continue
            q = nodes.index(n)
            x, y = int(s) % im.width, int(s) // im.width
            c = int(im.labels[s]) - 1
            for f in range(8):
                prm = hp.params(fd, 7, n, f, p.feat_low_bounds, p.feat_up_bounds, short)
                thr = hp.thresholds(4, 7, n, f, -tb, tb, short)
                v = hp.feature(kind, img, x, y, prm)
                for t in range(4):
                    want[q, c, t, f] += int(v <= thr[t])
    assert np.array_equal(H, want)


def _route_counts(ts, kind, out, depth):
    """leaf of every training sample through the oracle's tree (prediction path)."""
    tr = Tree(0, depth, {0: 0, 1: 1, 3: 3}[kind], ts.n_classes)
    tr.left_children[:] = out["left_children"]
    tr.features[:] = out["features"]
    tr.thresholds[:] = out["thresholds"]
    tr.posteriors[:] = out["posteriors"]
    counts = np.zeros((tr.n_nodes, ts.n_classes), dtype=np.int64)
    for im in ts.images:
        leaves = orc.predict_leaves(kind, im.data.reshape(im.height, im.width), None, tr).reshape(-1)
        for s in im.samples:
            counts[leaves[s], im.labels[s] - 1] += 1
    return tr, counts


def test_train_tree_invariants_c1_mini():
    ts = synth.training_set("c1", n_images=12, samples=128)
    p = synth.params_for("c1", n_features=40, n_thresholds=6)
    o = orc.Tset(0, ts)
    depth = 7
    out = orc.train_tree(o, p, ts.priors, 0, depth)
    lc = out["left_children"]
    n = lc.size
    assert lc[0] == 1, "root is always trained"
    for i in range(n):
        if lc[i] >= 0:
            assert lc[i] == 2 * i + 1 and lc[2 * i + 1] != -2 and lc[2 * i + 2] != -2
        if lc[i] == -2:
            assert not out["features"][i].any() and out["thresholds"][i] == 0
    tr, counts = _route_counts(ts, 0, out, depth)
    for i in range(n):
        if lc[i] == -1:
            # leaf posterior = class frequencies of the training samples reaching it
            c = counts[i]
            if i == 0:
                continue
            assert c.sum() > 0
            want = (c.astype(np.float32) / np.float32(c.sum())).astype(np.float32)
            assert out["posteriors"][i].tobytes() == want.tobytes(), i
            lvl = int(np.floor(np.log2(i + 1))) + 1
            # a leaf above the last level holds <= thr samples, or its best
            # split sent every sample one way (zero gain, _learn_best_featthr.hpp:145)
            if lvl < depth and c.sum() > p.per_leaf_samples_thr:
                assert abs(out["node_gain64"][i]) <= 1e-12, i
    # FP64 gain of each split equals the closed form of its children counts
    for i in range(n):
        if lc[i] >= 0:
            left = _subtree_counts(counts, lc, 2 * i + 1)
            right = _subtree_counts(counts, lc, 2 * i + 2)
            assert abs(out["node_gain64"][i] - hp.gain64(left, left + right)) < 1e-9
            assert left.sum() > 0 and right.sum() > 0


def _subtree_counts(counts, lc, i):
    if i >= lc.size or lc[i] == -2:
        return np.zeros(counts.shape[1], dtype=np.int64)
    if lc[i] == -1:
        return counts[i]
    return _subtree_counts(counts, lc, 2 * i + 1) + _subtree_counts(counts, lc, 2 * i + 2)


def test_small_node_not_split_and_root_gate():
    ts = synth.training_set("c1", n_images=2, samples=8)
    p = synth.params_for("c1", n_features=16, n_thresholds=4)
    p.per_leaf_samples_thr = 100    # more than the 16 samples: the root histogram is empty
    o = orc.Tset(0, ts)
    out = orc.train_tree(o, p, ts.priors, 0, 5)
    assert out["left_children"][0] == -2 and (out["left_children"][1:] == -2).all()
    assert out["posteriors"][0].tobytes() == ts.priors.tobytes()


def test_priors_running_mean():
    ts = synth.training_set("c1", n_images=7, samples=50)
    o = orc.Tset(0, ts)
    assert orc.tset_priors(o).tobytes() == ts.priors.tobytes()
    # the running mean stays within float rounding of the plain mean
    assert np.allclose(ts.priors, np.mean([im.priors for im in ts.images], axis=0), atol=1e-6)


def test_predict_brute_force():
    rng = np.random.default_rng(30)
    ts = synth.training_set("c1", n_images=6, samples=100)
    p = synth.params_for("c1", n_features=20, n_thresholds=5)
    o = orc.Tset(0, ts)
    trees = []
    for tid in range(3):
        out = orc.train_tree(o, p, ts.priors, tid, 6)
        t = Tree(tid, 6, 0, 4)
        t.left_children[:] = out["left_children"]
        t.features[:] = out["features"]
        t.thresholds[:] = out["thresholds"]
        t.posteriors[:] = out["posteriors"]
        trees.append(t)
    img, _ = synth.heldout("c1", n=1, w=40, h=30)[0]
    mask = (rng.uniform(size=img.shape) > 0.2).astype(np.uint8)
    post = orc.predict_posterior(0, img, mask, trees, 4)
    acc = np.zeros((4,) + img.shape, dtype=np.float32)
    for t in trees:
        leaves = orc.predict_leaves(0, img, mask, t)
        for y in range(img.shape[0]):
            for x in range(img.shape[1]):
                if not mask[y, x]:
                    assert leaves[y, x] == -1
                    continue
                n = 0
                while t.left_children[n] >= 0:
                    v = hp.feature(0, img, x, y, t.features[n])
                    n = t.left_children[n] + (0 if v <= t.thresholds[n] else 1)
                assert leaves[y, x] == n
                acc[:, y, x] += t.posteriors[n]
    want = (acc / np.float32(3)).astype(np.float32)
    assert post.tobytes() == want.tobytes()
    p64 = np.zeros(acc.shape)
    for t in trees:
        lv = orc.predict_leaves(0, img, mask, t)
        p64 += np.where(mask[None] > 0, t.posteriors[np.maximum(lv, 0)].transpose(2, 0, 1).astype(np.float64), 0)
    assert np.allclose(post, p64 / 3, rtol=1e-6, atol=1e-7)
