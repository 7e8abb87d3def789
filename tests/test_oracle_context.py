"""Pins of the oracle's auto-context restatement, on the CPU: the hook
against numpy brute force for given trees and node-id images, and the
routing of the node-id images (every routed pixel at a node of the right
depth, children of its previous node; unrouted pixels at -1)."""
import numpy as np

import padenti_amd as pa
from oracle import oracle as orc
from padenti_amd import synth


def test_context_hook_brute_force():
    rng = np.random.default_rng(0)
    h, w, nC = 11, 13, 3
    img = rng.random((h, w)).astype(np.float32) * 4
    post = rng.random((15, nC)).astype(np.float32)
    left = np.full(15, -1, np.int32)
    for trial in range(400):
        mp = rng.integers(-1, 15, (h, w)).astype(np.int32)
        p = np.array([rng.integers(-5, 6), rng.integers(-5, 6), rng.integers(0, 9), rng.random() * 2], np.float32)
        x, y = int(rng.integers(0, w)), int(rng.integers(0, h))
        ox, oy = x + int(p[0]), y + int(p[1])
        inside = 0 <= ox < w and 0 <= oy < h
        a = img[oy, ox] if inside else np.float32(10)
        n = mp[oy, ox] if inside else -1
        q = post[n, int(np.uint32(p[2])) % nC] if n >= 0 else np.float32(0)
        want = np.float32(np.float32(a - img[y, x]) + np.float32(p[3] * q))
        got = orc.feature_ac_at(img, x, y, p, left, post, mp)
        assert got == want, trial


def test_context_tree_routing_invariants():
    ts = synth.training_set("c1", n_images=4, samples=96)
    nC = 4
    p = pa.TreeTrainerParameters(48, 8, [-8, -8, 0, 0], [8, 8, 16, 2], -1.0, 1.0, 4)
    out = orc.train_tree(orc.Tset(orc.KIND_AC_DEPTH, ts), p, ts.priors, 0, 6)
    left = out["left_children"]
    assert (left >= 0).sum() >= 3

    class T:
        pass
    tr = T()
    tr.left_children, tr.features, tr.thresholds, tr.posteriors, tr.depth = (left, out["features"],
                                                                              out["thresholds"],
                                                                              out["posteriors"], 6)
    img = ts.images[0].data.reshape(120, 160)
    mask = (np.arange(160 * 120) % 7 != 0).astype(np.uint8).reshape(120, 160)
    leaves, post = orc.predict_ac(orc.KIND_AC_DEPTH, img, mask, [tr], nC)
    lv = leaves[0]
    assert np.all(lv[mask == 0] == -1) and np.all(lv[mask == 1] >= 0)
    for n in np.unique(lv[mask == 1]):
        assert left[n] < 0 and left[n] != -2   # every routed pixel ends at a leaf of the tree
    assert np.allclose(post[:, mask == 1].sum(0), 1.0, atol=1e-5)
    assert np.all(post[:, mask == 0] == 0)
