"""Auto-context hooks on the GPU: a hook that reads the tree being trained
and the node-id image (padenti_amd/examples/context_plugin.hip) trains trees
bit-identical to the oracle, which re-routes every labelled pixel from the
root at every level as the reference does (the device routes one level per
level), and predicts identical leaves and posteriors (masked pixels
included)."""
import os

import numpy as np
import pytest

import padenti_amd as pa
from oracle import oracle as orc
from padenti_amd import synth

pytestmark = pytest.mark.gpu
HERE = os.path.dirname(os.path.abspath(pa.__file__))


def context_plugin():
    return pa.Plugin(os.path.join(HERE, "libpd_context_plugin.so"), "depth_context", np.float32, 1, np.float32, 4)


@pytest.mark.parametrize("mode", [pa.PD_MODE_LEVEL_BATCHED, pa.PD_MODE_PER_NODE])
def test_context_hook_trains_and_predicts_like_the_oracle(gpu, mode):
    pl = context_plugin()
    ts = synth.training_set("c1", n_images=8, samples=160)
    nC = 4
    p = pa.TreeTrainerParameters(64, 10, [-8, -8, 0, 0], [8, 8, 16, 2], -1.0, 1.0, 4)
    trees = []
    tr = pa.HIPTreeTrainer(None, nC, plugin=pl)
    for tid, depth in ((0, 8), (1, 6)):
        tree = pa.Tree(tid, depth, pl, nC)
        tr.train(tree, ts, p, 1, depth, mode=mode)
        out = orc.train_tree(orc.Tset(orc.KIND_AC_DEPTH, ts), p, ts.priors, tid, depth)
        assert np.array_equal(tree.left_children, out["left_children"]), tid
        assert tree.features.tobytes() == out["features"].tobytes()
        assert tree.thresholds.tobytes() == out["thresholds"].tobytes()
        assert tree.posteriors.tobytes() == out["posteriors"].tobytes()
        assert (tree.left_children >= 0).sum() >= 7
        trees.append(tree)
    tr.close()
    # the context changes the trees: weight 0 (no posterior term) trains another tree
    p0 = pa.TreeTrainerParameters(64, 10, [-8, -8, 0, 0], [8, 8, 16, 0], -1.0, 1.0, 4)
    out0 = orc.train_tree(orc.Tset(orc.KIND_AC_DEPTH, ts), p0, ts.priors, 0, 8)
    assert not np.array_equal(out0["thresholds"], trees[0].thresholds)
    clf = pa.HIPClassifier(None, nC, plugin=pl)
    for t in trees:
        clf << t
    img, _ = synth.heldout("c1", n=1)[0]
    mask = (np.arange(160 * 120) % 5 != 0).astype(np.uint8).reshape(120, 160)
    leaves, post = orc.predict_ac(orc.KIND_AC_DEPTH, img, mask, trees, nC)
    for t in range(2):
        assert np.array_equal(clf.predict_leaves(t, img, 160, 120, mask), leaves[t]), t
    assert clf.predict_posterior(img, 160, 120, mask).tobytes() == post.tobytes()
    _, post_all = orc.predict_ac(orc.KIND_AC_DEPTH, img, None, trees, nC)
    assert clf.predict_posterior(img, 160, 120).tobytes() == post_all.tobytes()
    clf.close()
