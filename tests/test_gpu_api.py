"""The C++ drop-in API (tests/cpp/api_demo) and the feature-plugin path on
the GPU, checked against the CPU oracle."""
import os
import subprocess

import numpy as np
import pytest

import padenti_amd as pa
from padenti_amd import synth, treeio
from oracle import oracle as orc

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_cpp_api_demo_matches_oracle(gpu, tmp_path):
    from padenti_amd import build as b
    r = subprocess.run([b.DEMO_BIN, str(tmp_path)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    # the same training set through the oracle
    ts = synth.training_set("c1", n_images=8, samples=128)
    p = synth.params_for("c1", n_features=64, n_thresholds=10)
    out = orc.train_tree(orc.Tset(0, ts), p, ts.priors, 0, 6)
    rd = lambda n, dt: np.fromfile(str(tmp_path / n), dtype=dt)
    assert rd("priors.bin", np.float32).tobytes() == ts.priors.tobytes()
    assert np.array_equal(rd("left.bin", np.int32), out["left_children"])
    assert rd("features.bin", np.float32).tobytes() == out["features"].tobytes()
    assert rd("thresholds.bin", np.float32).tobytes() == out["thresholds"].tobytes()
    assert rd("posteriors.bin", np.float32).tobytes() == out["posteriors"].tobytes()
    tree = treeio.load(str(tmp_path / "tree.xml"), 0, 4)
    assert np.array_equal(tree.left_children, out["left_children"])
    assert tree.features.tobytes() == out["features"].tobytes()
    img = rd("image.bin", np.float32).reshape(120, 160)
    want = orc.predict_posterior(0, img, None, [tree], 4)
    assert rd("post.bin", np.float32).tobytes() == want.tobytes()
    mask = (np.arange(160 * 120) % 3 != 0).astype(np.uint8).reshape(120, 160)
    assert rd("post_masked.bin", np.float32).tobytes() == orc.predict_posterior(0, img, mask, [tree], 4).tobytes()
    assert np.array_equal(rd("leaves.bin", np.int32).reshape(120, 160), orc.predict_leaves(0, img, None, tree))
    # the directory-built training sets (uniform and class-weighted samplers)
    from oracle import ingest as oi
    counts, allv = rd("ds_counts.bin", np.uint32), rd("ds_samples.bin", np.uint32)
    o, acc = 0, np.zeros(8)
    for i in range(5):
        depth, lab = synth.gen_image(pa._lib.PD_SYN_CAPSULES_U16, 0, i, 96, 64, 8)
        want = [oi.sample_uniform(depth, lab, 300, seed=7), oi.sample_weighted(depth, lab, 300, 8, seed=7)]
        for k in range(2):
            assert np.array_equal(allv[o:o + counts[2 * i + k]], want[k]), (i, k)
            o += counts[2 * i + k]
        acc += pa.TrainingSetImage(depth, 96, 64, lab, 8, want[0]).priors
    assert rd("ds_priors.bin", np.float32).tobytes() == (acc / 5).astype(np.float32).tobytes()


def test_plugin_hook_trains_oracle_trees(gpu):
    plugin = pa.example_plugin()
    ts = synth.training_set("c1", n_images=10, samples=200)
    p = synth.params_for("c1", n_features=96, n_thresholds=10)
    tree = pa.Tree(2, 7, pa.PD_FEAT_DEPTH_F32, 4)
    tr = pa.HIPTreeTrainer(pa.PD_FEAT_DEPTH_F32, 4, plugin=plugin)
    tr.train(tree, ts, p, 1, 7)
    out = orc.train_tree(orc.Tset(0, ts), p, ts.priors, 2, 7)
    assert np.array_equal(tree.left_children, out["left_children"])
    assert tree.features.tobytes() == out["features"].tobytes()
    assert tree.posteriors.tobytes() == out["posteriors"].tobytes()
    clf = pa.HIPClassifier(pa.PD_FEAT_DEPTH_F32, 4, plugin=plugin)
    clf << tree
    img, _ = synth.heldout("c1", n=1)[0]
    assert clf.predict_posterior(img, 160, 120).tobytes() == orc.predict_posterior(0, img, None, [tree], 4).tobytes()
    with pytest.raises(pa.PdError):
        pa.HIPTreeTrainer(pa.PD_FEAT_DEPTH_F32, 0, plugin=plugin)
    tr.close()
    clf.close()
